// libmfa_hip device side: state, cell list, neighbour list, force pass and
// integrator for MI355X (gfx950), plus the C ABI declared in include/mfa_hip.h.
//
// Layout in HBM (one context = one system):
//   r, v, a, f     double[N][3]   the reference's (3,N) arrays, AoS, so a
//                                 neighbour gather is one 24-byte record
//   type           int32[N]       a_type (1-based)
//   mic            int32[N][3]    minimum-image counters (mic_count)
//   r_ref          double[N][3]   r0_old + mic_old*L at the last rebuild
//   cell_of        int32[N]       linear cell id (x fastest)
//   cell_count     int32[C]       particles per cell (zeroed by every integrate
//                                 kernel, filled by a rebuild)
//   cell_slots     int32[C][ccap] particle ids of a cell in arrival order; the
//                                 list build orders each cell by id in LDS (the
//                                 reference's bin_fluid order)
//   nl, nn         uint32[N][cap], int32[N]  full Verlet list (both directions);
//                                 an entry is j | code << 27, code (0..26) the
//                                 offset (dx,dy,dz) in {-1,0,1}^3 of j's cell from
//                                 i's cell at the rebuild, 9(dx+1)+3(dy+1)+(dz+1)
//   pos4, vel4     float4[N]      FP32 view of the mixed form: position relative
//                                 to the origin of the particle's cell at the
//                                 last rebuild (.w = type-1 as int bits), and
//                                 velocity; rewritten whenever r or v change
//   step           int64          current step (RNG key, switch-on ramp); lives
//                                 on the device so captured graphs replay
//
// Everything in this file is compiled with -ffp-contract=off: the FP64 forms
// reproduce the reference's operation order; the FP32 helpers of the mixed
// form enable contraction locally.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/mfa_hip.h"
#include "mfa_rng.h"

#define MFA_BLOCK 256
#define MFA_NTERMS 12
// term slots of the per-block partial sums
enum {
  T_VFF = 0, T_VFW, T_VINTRA, T_VBEND, T_VCOUL, T_TKIN, T_VIR, T_FZB, T_FZT, T_ZMOM, T_SPARE0, T_SPARE1
};
// force-pass term mask
enum { TM_PAIR = 1, TM_WALL = 2, TM_BOND = 4, TM_ACCUM = 8, TM_ENERGY = 16, TM_BEND = 32, TM_FUSE = 64 };
#define TM_ALL (TM_PAIR | TM_WALL | TM_BOND | TM_BEND)
// device flag slots
// F_MIRROR: the list was rebuilt since the half-list pair pass last found
// the mirror positions of its lower entries
enum { F_REBUILD = 0, F_ERROR = 1, F_MAXNEIGH = 2, F_NREBUILD = 3, F_FILLDONE = 4, F_LIVE = 5,
       F_MIRROR = 6, F_NFLAGS = 8 };

// Fixed-point scale of the mixed form: 2^22 per unit force (resolution
// 2.4e-7; a pair term converts through int32 up to |F| < 512, through int64
// beyond; both truncate, so the two routes agree).
#define MFA_FIX_SCALE 4194304.0f
#define MFA_FIX_INV (1.0 / 4194304.0)
#define MFA_JBITS 27
#define MFA_JMASK ((1u << MFA_JBITS) - 1u)
// List rows.  Row s (sorted position s) holds its "upper" entries (neighbours
// at sorted positions above s) at [0, nup) in scan order and its "lower"
// entries at [cap - nlo, cap), appended from the back (reverse scan order);
// nn[s] = (nup << 16) | (nup + nlo).  A kernel that walks the whole row maps
// its q-th entry to list_slot(q, nup, cnt, cap).
#define MFA_CAP_MAX 32767
// Row layouts (what the front part holds): everything; the upper entries
// (half-list form); the entries inside the interaction range at the build
// (near/shell rows, the skin part in the back)
enum { LK_PLAIN = 0, LK_HALF = 1, LK_NEAR = 2 };
__device__ __forceinline__ int list_slot(int q, int nup, int cnt, int cap) {
  return q < nup ? q : q + (cap - cnt);
}

struct KParams {
  double L[3], invL[3];
  double dt, dt_2, half_dt, half_dt_2;
  double inv_mass_type[MFA_NTYPE], mass_type[MFA_NTYPE];
  double epsil[MFA_NTYPE][MFA_NTYPE], sigma_2[MFA_NTYPE][MFA_NTYPE];
  double range_2[MFA_NTYPE][MFA_NTYPE], inv_range_2[MFA_NTYPE][MFA_NTYPE];
  double e_shift[MFA_NTYPE][MFA_NTYPE], list_2[MFA_NTYPE][MFA_NTYPE];
  double r_cut_dpd_2, friction1, sig;
  // 1 / range and 1 / r_cut_dpd (the DPD weight w_r = 1 - r / r_c)
  double inv_range[MFA_NTYPE][MFA_NTYPE], inv_r_cut_dpd;
  double k_chain, inv_r_chain_2;
  double a_w, sigma_wall[MFA_NTYPE], a_wall[MFA_NTYPE], z_space_wall;
  double va[3];
  double inv_r_bin[3], r_bin[3];
  double skin_half_2;
  double skin_2;  // list_2 - skin_2: the interaction range^2 of a type pair
  double fso_r2_init, fso_time;  // force switch-on ramp (0 = off)
  // Coulomb (SYSTEM 2/3): charge per type, erfc(sqrt(alpha_c) r) real space
  // with the prefactor inv_sqrt_pi (1/pi in the compat form,
  // init_config.f90:230), inside rc_coul or (rc_coul_2 = 0) inside the LJ range
  double q_type[MFA_NTYPE], alpha_c, sqrt_alpha, inv_sqrt_pi, rc_coul_2;
  int coulomb;
  int bending;                // BENDING: harmonic angles of the brush chains
  double k_bend, alpha_eq;
  float pf_margin;               // bound on |r^2(FP32) - r^2(FP64)| in the list build
  float ff_margin;               // the same for the cell-relative FP32 view of the force pass
  float lf_margin;               // ... and of the list build, against the list radius
  int nb[3];
  int n_cells, cell_cap;
  // replica layout: particles [p0, p0 + n) (p0 a multiple of 256), cells
  // [c0, c0 + n_cells) of the concatenated arrays; flags at rep * F_NFLAGS
  int p0, c0, rep;
  int fg0;  // first of the replica's completion counters of the cell fill
  int n, n_mon, n_chain, n_brush, part_init_d, n_mon_d, part_init_e, n_mon_e;
  int symmetry, pinned_heads, dpd_on, dpd_cut_off_on, dpd_head_excl;
  int dpd_vv;  // DPD_VV: heads move, new_dpd_fd after verlet_velocities
  int lgv;     // THERMOSTAT 1: Langevin force on every particle instead of DPD
  uint32_t seed;
};

// Ewald k-space launch constants (the k grid lives in device arrays).
struct EwaldK {
  double dk[3];       // delta_k (ewald_k.f90:44-52)
  double lz_half;     // init_config.f90:246
  int nk, nsegs, nq;
  int nxm, nym, nzm;  // largest |nx|, |ny|, nz of the k grid
  int nex, ney, nez;  // phase-table entries per particle and axis
  int n_chunks;       // structure-factor chunks of EW_CH charged particles
  int n_mz;           // dipole-moment partials (table blocks)
  int segs_per_kgroup;  // forces: segments per block row
};

// Host plan of the k-space stage: the k grid of a box (ewald_grid) and the
// launch shapes for a number of charged particles (ewald_shape), checked
// against every bound the kernels rely on before anything is launched.
struct EwPlan {
  EwaldK e;
  double inv_V;
  int sum_blocks;                   // k_ewald_sum blocks (64 k vectors each)
  std::vector<double4> kv;          // (kx, ky, kz, exp(-k^2/(4 alpha_c))/k^2)
  std::vector<int4> segs, rows;     // segments (ny, nz, nx_lo, len), rows (ny, nz, nx_lo, nx_hi)
  std::vector<int> seg_k0, row_k0;  // k index of each segment's / row's first vector
  bool rows_kernel;                 // forces row by row (k_ewald_force_rows)
  int rows_per_group, n_kgroups;    // force partial groups (row groups or segment groups)
  int force_blocks, force_grid;     // row kernel: blocks, and the grid padded to a multiple of 8
  size_t lds_rho, lds_force;        // dynamic LDS bytes of k_ewald_rho and the force kernel
  size_t fk_cap;                    // force partials (per particle and group) the buffer holds
};

struct mfa_ctx {
  mfa_params p;  // replica 0 (all replicas share the force-field switches)
  // replicas: particles [p0, p0 + n) and cells [c0, c0 + n_cells) of the
  // concatenated arrays; n and n_cells below are the padded totals
  int n_rep;
  std::vector<mfa_params> ps;
  std::vector<KParams> ks;
  KParams* d_kp;     // per replica
  int* d_rmap;       // replica of each 256-particle block
  int* d_cell_rep;   // replica of each cell
  std::vector<mfa_obs> rep_obs;  // per-replica observations of the last mfa_run_steps
  int precision;
  bool extra_terms;  // some replica has charges or bending: force kernels with XT & 1
  bool lgv_terms;    // some replica runs THERMOSTAT 1: force kernels with XT & 2
  int n;
  int cap;
  bool cap_fixed;
  // mixed pass, how list entries are walked: 0 directly, 1 through the pair
  // queue (the interacting pairs of a wave compacted), 2 the near part
  // directly and the skin part through the queue (near/shell rows)
  int pair_queue;
  // half-list pair passes of the mixed form: each pair evaluated once, from
  // the row of its lower sorted position; its scaled pair scalar kept per
  // upper entry (pair_s), the position of each lower entry's pair in its
  // neighbour's row (mir), and the first pass's fixed-point sums (fpart)
  bool half_list;
  int list_kind;  // row layout the builds make: LK_PLAIN, LK_HALF or LK_NEAR
  float* pair_s = nullptr;
  uint16_t* mir = nullptr;
  double* fpart = nullptr;
  int half_cap = 0;  // the list capacity those buffers were sized for
  int tpp;  // threads per particle in the force pass
  int n_cells;
  bool need_rebuild;
  hipStream_t stream;
  double *r, *v, *a, *f, *r_ref;
  int32_t *type, *mic;
  int32_t *cell_of, *cell_count, *cell_slots, *cell_start;
  int32_t *perm, *sidx;  // sorted position -> particle, particle -> sorted position
  int32_t* fill_cnt;     // completion counters of the cell fill (zero between launches)
  bool lb_octants;       // list build with octant groups (dense cells)
  // observation stage (mfa_obs_init): head forces of the last
  // verlet_velocities, slab counts, per-chain Rg^2, device sums
  double* head_f;
  unsigned* d_hist;
  double *d_chain_rg, *d_obs;
  int hist_dim;
  bool obs_on;
  mfa_obs_sums obs_host;  // the energy sums, kept on the host
  int cell_cap;
  unsigned* nl;
  int32_t* nn;
  float4 *pos4, *vel4;  // FP32 view of the mixed form
  // the other buffers of positions, velocities and view: a step whose force
  // pass also integrates reads r, v, pos4, vel4 and writes these (the FP64
  // pass reads the velocities of its neighbours, which other blocks of the
  // same pass integrate); the host then swaps the pointers (parity counts
  // the swaps)
  double *r_alt, *v_alt;
  float4 *pos4_alt, *vel4_alt;
  int parity;
  // set only while mfa_time_force_step times that force pass: copies that
  // take what it writes to v, mic, the flags and the cell counts
  double* probe_v = nullptr;
  int32_t *probe_mic = nullptr, *probe_flags = nullptr, *probe_count = nullptr;
  double *fd, *fd_new;  // DPD_VV: dissipative force of the last new_dpd_fd, and the new one
  bool dpd_vv;
  // Ewald k space (charged systems): the checked host plan and its tables
  bool ewald;
  EwPlan ew;
  bool ew_rows_ok;  // the row kernel exists for the grid and does not spill
  size_t lds_limit; // LDS bytes per workgroup the kernels may use
  double4* d_kvec;
  int4* d_runs;     // segments (ny, nz, nx_lo, length), longest first
  int* d_run_k0;    // k index of each segment's first vector
  int4* d_rows;     // rows: maximal runs (ny, nz, nx_lo, nx_hi) of consecutive kx
  int* d_row_k0;    // k index of each row's first vector
  double2 *d_tx, *d_ty, *d_tz;  // phase tables [entry][charged particle]
  int* d_qidx;
  double* d_qval;
  int* d_qinv;      // charged index of each particle, -1 for the neutral ones
  double2 *d_rho_part, *d_grho;
  double *d_uk_part, *d_mz_part, *d_mz, *d_fk_part, *d_ek;
  double v_coul_self;
  long long* d_step;    // device step counter
  // captured step block: graph_steps steps of [rebuild, force, integrate]
  int graph_steps;
  hipGraphExec_t gexec;
  int gexec_precision, gexec_tpp, gexec_parity;
  bool gexec_fused;
  bool capturing;
  int32_t* flags;       // device flags
  int32_t* h_flags;     // pinned mirror
  double* part;         // per-block partial sums [blocks][MFA_NTERMS]
  double* red;          // final sums [MFA_NTERMS]
  double* h_red;        // pinned mirror
  int max_blocks;
  // timing: pairs of events recorded around each timed launch, read back
  // only when the caller asks (no per-launch synchronisation)
  bool timing;
  std::vector<hipEvent_t> ev_force, ev_integ;
  size_t ev_force_used, ev_integ_used;
  double force_ms;
  int64_t force_n;
  double integ_ms;
  int64_t integ_n;
  std::string err;
};

static thread_local std::string g_noctx_err;

extern "C" void mfa_internal_set_noctx_error(const char* msg) { g_noctx_err = msg; }

static int set_err(mfa_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  else g_noctx_err = buf;
  return code;
}

#define HIPCHK(c, call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return set_err((c), MFA_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));    \
  } while (0)

// ---------------------------------------------------------------------------
// Device helpers

// PBC as the reference writes it: d - L*int(2*d/L), int() truncating
// (binning.f90:89-93, fluid_fluid.f90:84-88, intra_molec.f90:25-29).
__device__ __forceinline__ double pbc(double d, double L, double invL) {
  const int m = (int)((2.0 * d) * invL);
  return d - L * (double)m;
}

__device__ __forceinline__ void min_image(const KParams& k, double& dx, double& dy, double& dz) {
  dx = pbc(dx, k.L[0], k.invL[0]);
  dy = pbc(dy, k.L[1], k.invL[1]);
  if (k.symmetry == 1) dz = pbc(dz, k.L[2], k.invL[2]);
}

__device__ __forceinline__ int cell_coord(double x, double inv_r_bin, int nb) {
  // mod(int(r*inv_r_bin), n_bin) (binning.f90:27-29); Fortran mod == C %.
  return ((int)(x * inv_r_bin)) % nb;
}

__device__ __forceinline__ bool is_head(const KParams& k, int i) {
  return k.pinned_heads && i < k.n_brush && (i % k.n_mon) == 0;
}

// x**9 and friends as the integer-power chains a compiler emits.
__device__ __forceinline__ double pow3(double x) { return (x * x) * x; }

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Block-level sum of MFA_NTERMS doubles held per thread; thread 0 writes the
// block's partials.  Deterministic (fixed tree).
__device__ void block_partials(double* vals, double* part_out) {
  __shared__ double sh[MFA_BLOCK / 64][MFA_NTERMS];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int t = 0; t < MFA_NTERMS; ++t) {
    double v = vals[t];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) sh[wid][t] = v;
  }
  __syncthreads();
  if (threadIdx.x < MFA_NTERMS) {
    double s = 0.0;
    for (int w = 0; w < MFA_BLOCK / 64; ++w) s += sh[w][threadIdx.x];
    part_out[(size_t)blockIdx.x * MFA_NTERMS + threadIdx.x] = s;
  }
}

// Final sums: one block per term; each thread adds a fixed strided subset
// of the block partials, then a fixed tree combines the threads.
__global__ __launch_bounds__(MFA_BLOCK) void k_reduce_partials(const double* __restrict__ part,
                                                               int b0, int nblocks,
                                                               double* __restrict__ out) {
  __shared__ double sh[MFA_BLOCK];
  const int t = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += MFA_BLOCK) s += part[(size_t)(b0 + b) * MFA_NTERMS + t];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = MFA_BLOCK / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[t] = sh[0];
}

// The same per replica: block (term, replica) sums the replica's partials,
// `per` particles per block (256 / threads per particle, or 256).
__global__ __launch_bounds__(MFA_BLOCK) void k_reduce_replicas(const double* __restrict__ part,
                                                               const KParams* __restrict__ kb,
                                                               int per, double* __restrict__ out) {
  __shared__ double sh[MFA_BLOCK];
  const int t = blockIdx.x, q = blockIdx.y;
  const int b0 = kb[q].p0 / per, nb = (kb[q].n + per - 1) / per;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += MFA_BLOCK) s += part[(size_t)(b0 + b) * MFA_NTERMS + t];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = MFA_BLOCK / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(size_t)q * MFA_NTERMS + t] = sh[0];
}

// ---------------------------------------------------------------------------
// Integrator kernels (FP64).

// verlet_positions (verlet_positions.f90:144-170, wrap 233-250).
__device__ __forceinline__ void vp_one(const KParams& k, double* r, double* v, const double* a,
                                       int* mic, bool& wrapped) {
  for (int d = 0; d < 3; ++d) {
    r[d] = (r[d] + k.dt * v[d]) + k.half_dt_2 * a[d];
    v[d] = v[d] + k.half_dt * a[d];
  }
  const int nd = k.symmetry == 0 ? 2 : 3;
  for (int d = 0; d < nd; ++d) {
    if (r[d] > k.L[d]) {
      r[d] = r[d] - k.L[d];
      mic[d] += 1;
      wrapped = true;
    } else if (r[d] < 0.0) {
      r[d] = r[d] + k.L[d];
      mic[d] -= 1;
      wrapped = true;
    }
  }
}

// check_skin for one particle (check_skin.f90:8-21).
__device__ __forceinline__ bool skin_one(const KParams& k, const double* r, const double* rref,
                                         const int* mic) {
  double s = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double dd = (rref[d] - r[d]) - (double)mic[d] * k.L[d];
    s = s + dd * dd;
  }
  return s > k.skin_half_2;
}

// verlet_velocities for one particle (verlet_velocities.f90:25-36,120-133).
__device__ __forceinline__ void vv_one(const KParams& k, int i, int ti, double* f, double* v,
                                       double* a) {
  const double im = k.inv_mass_type[ti];
  for (int d = 0; d < 3; ++d) {
    a[d] = f[d] * im;
    v[d] = v[d] + k.half_dt * a[d];
  }
  // with DPD_VV the head block of verlet_velocities is compiled out
  // (verlet_velocities.f90:79-230 is #ifndef DPD_VV)
  if (is_head(k, i) && !k.dpd_vv) {
    const int c = i / k.n_mon;
    const double sgn = c < k.n_chain / 2 ? 1.0 : -1.0;
    v[0] = sgn * k.va[0];
    v[1] = sgn * k.va[1];
    v[2] = 0.0;
    for (int d = 0; d < 3; ++d) {
      f[d] = 0.0;
      a[d] = 0.0;
    }
  }
}

// Clears the cell counts of every replica for the rebuild that may follow
// (grid-stride).
__device__ __forceinline__ void zero_counts(int nc_tot, int* __restrict__ cell_count) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc_tot; c += gridDim.x * blockDim.x)
    cell_count[c] = 0;
}

// Parameters of the replica a block works on.  Replicas own 256-aligned
// particle ranges of the concatenated arrays, so a block of particle-indexed
// work never spans two; a single system is a batch of one.
#define MFA_REPLICA(first_particle) const KParams& k = kb[rmap[(first_particle) >> 8]]

// FP32 view of particle i: position relative to the origin of `cell` (its
// cell at the last rebuild), wrapped back across periodic faces it crossed
// since, and velocity.
__device__ __forceinline__ void write_view(const KParams& k, int i, int cell, int ti, int il,
                                           const double* ri, const double* vi,
                                           float4* __restrict__ pos4, float4* __restrict__ vel4) {
  const int cx = cell % k.nb[0], cy = (cell / k.nb[0]) % k.nb[1], cz = cell / (k.nb[0] * k.nb[1]);
  double lx = ri[0] - (double)cx * k.r_bin[0];
  double ly = ri[1] - (double)cy * k.r_bin[1];
  double lz = ri[2] - (double)cz * k.r_bin[2];
  lx -= k.L[0] * rint(lx * k.invL[0]);
  ly -= k.L[1] * rint(ly * k.invL[1]);
  if (k.symmetry == 1) lz -= k.L[2] * rint(lz * k.invL[2]);
  pos4[i] = make_float4((float)lx, (float)ly, (float)lz, __int_as_float((il << 2) | ti));
  vel4[i] = make_float4((float)vi[0], (float)vi[1], (float)vi[2], 0.0f);
}

__global__ void k_verlet_positions(const KParams* __restrict__ kb, const int* __restrict__ rmap,
                                   double* __restrict__ r, double* __restrict__ v,
                                   const double* __restrict__ a, int* __restrict__ mic,
                                   const double* __restrict__ r_ref, int* __restrict__ flags,
                                   int with_skin, int nc_tot, int* __restrict__ cell_count,
                                   const int* __restrict__ type, const int* __restrict__ cell_of,
                                   const int* __restrict__ sidx, float4* __restrict__ pos4,
                                   float4* __restrict__ vel4) {
  zero_counts(nc_tot, cell_count);
  MFA_REPLICA(blockIdx.x * blockDim.x);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double over = 0.0;
  if (i - k.p0 < k.n) {
    double ri[3], vi[3], ai[3];
    int mi[3];
    for (int d = 0; d < 3; ++d) {
      ri[d] = r[3 * (size_t)i + d];
      vi[d] = v[3 * (size_t)i + d];
      ai[d] = a[3 * (size_t)i + d];
      mi[d] = mic[3 * (size_t)i + d];
    }
    bool wrapped = false;
    vp_one(k, ri, vi, ai, mi, wrapped);
    for (int d = 0; d < 3; ++d) {
      r[3 * (size_t)i + d] = ri[d];
      v[3 * (size_t)i + d] = vi[d];
    }
    if (wrapped)
      for (int d = 0; d < 3; ++d) mic[3 * (size_t)i + d] = mi[d];
    write_view(k, sidx[i], cell_of[i], type[i] - 1, i - k.p0, ri, vi, pos4, vel4);
    if (with_skin) {
      double rr[3];
      for (int d = 0; d < 3; ++d) rr[d] = r_ref[3 * (size_t)i + d];
      over = skin_one(k, ri, rr, mi) ? 1.0 : 0.0;
    }
  }
  if (with_skin) {
    over = wave_max(over);
    if ((threadIdx.x & 63) == 0 && over > 0.0) atomicOr(&flags[k.rep * F_NFLAGS + F_REBUILD], 1);
  }
}

__global__ void k_check_skin(const KParams* __restrict__ kb, const int* __restrict__ rmap,
                             const double* __restrict__ r, const double* __restrict__ r_ref,
                             const int* __restrict__ mic, int* __restrict__ flags) {
  MFA_REPLICA(blockIdx.x * blockDim.x);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double over = 0.0;
  if (i - k.p0 < k.n) {
    double ri[3], rr[3];
    int mi[3];
    for (int d = 0; d < 3; ++d) {
      ri[d] = r[3 * (size_t)i + d];
      rr[d] = r_ref[3 * (size_t)i + d];
      mi[d] = mic[3 * (size_t)i + d];
    }
    over = skin_one(k, ri, rr, mi) ? 1.0 : 0.0;
  }
  over = wave_max(over);
  if ((threadIdx.x & 63) == 0 && over > 0.0) atomicOr(&flags[k.rep * F_NFLAGS + F_REBUILD], 1);
}

// verlet_velocities alone: writes v, a and zeroes head forces; kinetic energy
// partials when `energy` (verlet_velocities.f90:155-177).  Particle i of the
// block's replica k; the FP32 view (vel4 at sidx) and the head forces are
// written when their arrays are given.
__device__ __forceinline__ void vv_body(const KParams& k, int i, double* __restrict__ f,
                                        double* __restrict__ v, double* __restrict__ a,
                                        const int* __restrict__ type, double* __restrict__ part,
                                        int energy, const int* __restrict__ sidx,
                                        float4* __restrict__ vel4, double* __restrict__ head_f) {
  double vals[MFA_NTERMS];
  for (int t = 0; t < MFA_NTERMS; ++t) vals[t] = 0.0;
  if (i - k.p0 < k.n) {
    const int ti = type[i] - 1;
    double fi[3], vi[3], ai[3];
    for (int d = 0; d < 3; ++d) {
      fi[d] = f[3 * (size_t)i + d];
      vi[d] = v[3 * (size_t)i + d];
    }
    // f_on_heads = force(:, heads) before the heads are pinned
    // (verlet_velocities.f90:14-19); single-system contexts
    if (head_f && is_head(k, i - k.p0))
      for (int d = 0; d < 3; ++d) head_f[3 * (size_t)((i - k.p0) / k.n_mon) + d] = fi[d];
    vv_one(k, i - k.p0, ti, fi, vi, ai);
    for (int d = 0; d < 3; ++d) {
      v[3 * (size_t)i + d] = vi[d];
      a[3 * (size_t)i + d] = ai[d];
    }
    if (vel4) vel4[sidx[i]] = make_float4((float)vi[0], (float)vi[1], (float)vi[2], 0.0f);
    if (is_head(k, i - k.p0) && !k.dpd_vv)
      for (int d = 0; d < 3; ++d) f[3 * (size_t)i + d] = 0.0;
    vals[T_TKIN] = k.mass_type[ti] * ((vi[0] * vi[0] + vi[1] * vi[1]) + vi[2] * vi[2]);
  }
  if (energy) block_partials(vals, part);
}

__global__ void k_verlet_velocities(const KParams* __restrict__ kb, const int* __restrict__ rmap,
                                    double* __restrict__ f, double* __restrict__ v,
                                    double* __restrict__ a, const int* __restrict__ type,
                                    double* __restrict__ part, int energy,
                                    const int* __restrict__ sidx, float4* __restrict__ vel4,
                                    double* __restrict__ head_f) {
  MFA_REPLICA(blockIdx.x * blockDim.x);
  vv_body(k, blockIdx.x * blockDim.x + threadIdx.x, f, v, a, type, part, energy, sidx, vel4, head_f);
}

// The same for one system on arrays the caller owns, its constants passed by
// value (no context, no replica table).
__global__ void k_verlet_velocities_1(KParams k, double* __restrict__ f, double* __restrict__ v,
                                      double* __restrict__ a, const int* __restrict__ type,
                                      double* __restrict__ part, int energy) {
  vv_body(k, blockIdx.x * blockDim.x + threadIdx.x, f, v, a, type, part, energy, nullptr, nullptr,
          nullptr);
}

// Fused: verlet_velocities of step s, then verlet_positions and check_skin of
// step s+1.  Reads F, v, r, mic, r_ref; writes v, r (mic only on a wrap).
__global__ void k_fused_integrate(const KParams* __restrict__ kb, const int* __restrict__ rmap,
                                  const double* __restrict__ f, double* __restrict__ v,
                                  double* __restrict__ r, int* __restrict__ mic,
                                  const double* __restrict__ r_ref, const int* __restrict__ type,
                                  int* __restrict__ flags, double* __restrict__ part, int energy,
                                  int nc_tot, int* __restrict__ cell_count,
                                  long long* __restrict__ d_step, const int* __restrict__ cell_of,
                                  const int* __restrict__ sidx, float4* __restrict__ pos4,
                                  float4* __restrict__ vel4) {
  zero_counts(nc_tot, cell_count);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *d_step += 1;  // next step; read by the next force pass only
  MFA_REPLICA(blockIdx.x * blockDim.x);
  if (flags[k.rep * F_NFLAGS + F_ERROR] != 0) return;
  double over = 0.0;
  double vals[MFA_NTERMS];
  for (int t = 0; t < MFA_NTERMS; ++t) vals[t] = 0.0;
  if (i - k.p0 < k.n) {
    const int ti = type[i] - 1;
    double fi[3], vi[3], ai[3], ri[3], rr[3];
    int mi[3];
    for (int d = 0; d < 3; ++d) {
      fi[d] = f[3 * (size_t)i + d];
      vi[d] = v[3 * (size_t)i + d];
      ri[d] = r[3 * (size_t)i + d];
      mi[d] = mic[3 * (size_t)i + d];
      rr[d] = r_ref[3 * (size_t)i + d];
    }
    vv_one(k, i - k.p0, ti, fi, vi, ai);
    if (energy) vals[T_TKIN] = k.mass_type[ti] * ((vi[0] * vi[0] + vi[1] * vi[1]) + vi[2] * vi[2]);
    bool wrapped = false;
    vp_one(k, ri, vi, ai, mi, wrapped);
    for (int d = 0; d < 3; ++d) {
      r[3 * (size_t)i + d] = ri[d];
      v[3 * (size_t)i + d] = vi[d];
    }
    if (wrapped)
      for (int d = 0; d < 3; ++d) mic[3 * (size_t)i + d] = mi[d];
    write_view(k, sidx[i], cell_of[i], ti, i - k.p0, ri, vi, pos4, vel4);
    over = skin_one(k, ri, rr, mi) ? 1.0 : 0.0;
  }
  over = wave_max(over);
  if ((threadIdx.x & 63) == 0 && over > 0.0) atomicOr(&flags[k.rep * F_NFLAGS + F_REBUILD], 1);
  if (energy) block_partials(vals, part);
}

// ---------------------------------------------------------------------------
// Cell list (binning.f90:11-50) in fixed-capacity cells.  Each rebuild kernel
// returns at once unless the rebuild flag is set, so a step can enqueue the
// rebuild unconditionally and let the skin test decide on the device.

__global__ void k_zero_counts(int nc_tot, int* __restrict__ cell_count) {
  zero_counts(nc_tot, cell_count);
}

__global__ void k_set_step(long long* __restrict__ d_step, long long step) { *d_step = step; }

// Sets flag `idx` of every replica.
__global__ void k_set_flag(int* __restrict__ flags, int n_rep, int idx, int value) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n_rep; q += gridDim.x * blockDim.x)
    flags[q * F_NFLAGS + idx] = value;
}

// Cell of each particle, mod(int(r*inv_r_bin), n_bin) per axis, and a slot in
// that cell (binning.f90:25-48).  Lanes take the particles in the spatial
// order of the last rebuild (s -> perm[s]), so a wave's particles come in
// runs of equal cells: one atomic per run reserves its slots.
#define FILL_GROUP 16  // blocks per first-level completion counter
#define FILL_KEEP 32   // cells per thread whose counts the scan keeps in registers
__global__ __launch_bounds__(MFA_BLOCK) void k_cell_fill(
    const KParams* __restrict__ kb, const int* __restrict__ rmap, const double* __restrict__ r,
    const int* __restrict__ perm, int* __restrict__ cell_of, int* __restrict__ cell_count,
    int* __restrict__ cell_slots, int* __restrict__ cell_start, int* __restrict__ flags_all,
    int* __restrict__ fill_cnt, long long* __restrict__ d_step, int inc_step) {
  // a step whose force pass integrates as well moves the step counter on
  // here, ahead of that force pass (no kernel of the rebuild reads it)
  if (inc_step && blockIdx.x == 0 && threadIdx.x == 0) *d_step += 1;
  MFA_REPLICA(blockIdx.x * blockDim.x);
  int* flags = flags_all + k.rep * F_NFLAGS;
  // F_REBUILD (raised by the skin checks) is handed on here: the replica's
  // last block below moves it to F_LIVE, which the sort and list kernels
  // read, and clears it; without a rebuild, the replica's first block clears
  // F_LIVE.  So no kernel after the fill needs a completion count of its own.
  if (flags[F_REBUILD] == 0) {  // uniform over the replica and the launch
    if (threadIdx.x == 0 && (int)blockIdx.x == k.p0 / MFA_BLOCK) flags[F_LIVE] = 0;
    return;
  }
  const bool ok = flags[F_ERROR] == 0;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int cell = -1, i = 0;
  if (ok && s - k.p0 < k.n) {
    i = perm[s];
    const int cx = cell_coord(r[3 * (size_t)i + 0], k.inv_r_bin[0], k.nb[0]);
    const int cy = cell_coord(r[3 * (size_t)i + 1], k.inv_r_bin[1], k.nb[1]);
    const int cz = cell_coord(r[3 * (size_t)i + 2], k.inv_r_bin[2], k.nb[2]);
    if (cx < 0 || cy < 0 || cz < 0) {  // the reference would index out of bounds
      atomicCAS(&flags[F_ERROR], 0, MFA_ERR_CELL);
      cell_of[i] = 0;
    } else {
      cell = (cz * k.nb[1] + cy) * k.nb[0] + cx;  // within the replica's grid
      cell_of[i] = cell;
    }
  }
  // slots: each run of lanes with the same cell (the spatial order makes
  // runs long) reserves its slots with one atomic, all runs of the wave at once
  const int prev = __shfl_up(cell, 1, 64);
  const unsigned long long starts = __ballot(lane == 0 || cell != prev);
  const unsigned long long upto = starts & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
  const int r0 = 63 - __builtin_clzll(upto);  // first lane of this lane's run
  const unsigned long long above = starts & ~(lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
  const int r1 = above ? __builtin_ctzll(above) : 64;  // first lane of the next run
  int base = 0;
  if (cell >= 0 && lane == r0) base = atomicAdd(&cell_count[k.c0 + cell], r1 - r0);
  base = __shfl(base, r0, 64);
  if (cell >= 0) {
    const int slot = base + (lane - r0);
    if (slot < k.cell_cap)
      cell_slots[(size_t)(k.c0 + cell) * k.cell_cap + slot] = i;
    else
      atomicCAS(&flags[F_ERROR], 0, MFA_ERR_CELL_OVERFLOW);
  }
  // The replica's last block to finish scans its cell counts:
  // cell_start[c] = p0 + particles in the replica's cells before c.  No block
  // waits for another.  Completion is counted in two levels (blocks of a
  // group of FILL_GROUP, then groups), so no counter takes more than a few
  // dozen atomics; each counter is reset by the block that finds itself last.
  // The counts are only ever changed by device-scope atomics, each of which
  // has returned (its slot is used above) before the barrier; the last block
  // reads them with device-scope atomic loads.  No cache-wide fence is needed.
  __shared__ int last_s, wsum[MFA_BLOCK / 64], carry_s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nb = (k.n + MFA_BLOCK - 1) / MFA_BLOCK;
    const int b = (int)blockIdx.x - k.p0 / MFA_BLOCK;  // block within the replica
    const int g = b / FILL_GROUP, ng = (nb + FILL_GROUP - 1) / FILL_GROUP;
    const int gsz = min(FILL_GROUP, nb - g * FILL_GROUP);
    int last = 0;
    if (atomicAdd(&fill_cnt[k.fg0 + g], 1) == gsz - 1) {
      atomicExch(&fill_cnt[k.fg0 + g], 0);
      last = atomicAdd(&flags[F_FILLDONE], 1) == ng - 1;
    }
    last_s = last;
    carry_s = k.p0;
  }
  __syncthreads();
  if (!last_s) return;
  if (threadIdx.x == 0) {  // every block of the replica has read F_REBUILD by now
    atomicExch(&flags[F_FILLDONE], 0);
    atomicExch(&flags[F_LIVE], 1);
    atomicExch(&flags[F_REBUILD], 0);
  }
  // one pass: thread t sums the counts of cells [t per, (t+1) per), a block
  // scan of the 256 sums gives each thread its offset, then its cells' starts.
  // Up to FILL_KEEP cells per thread the counts are read once, all loads in
  // flight together, and kept in registers for the second pass.
  const int per = (k.n_cells + MFA_BLOCK - 1) / MFA_BLOCK;
  const int cb = threadIdx.x * per, ce = min(k.n_cells, cb + per);
  int tot = 0;
  int kept[FILL_KEEP];
  const bool keep = per <= FILL_KEEP;
  if (keep) {
#pragma unroll
    for (int u = 0; u < FILL_KEEP; ++u)
      kept[u] = cb + u < ce ? __hip_atomic_load(&cell_count[k.c0 + cb + u], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT)
                            : 0;
#pragma unroll
    for (int u = 0; u < FILL_KEEP; ++u) {
      kept[u] = min(kept[u], k.cell_cap);
      tot += kept[u];
    }
  } else {
#pragma unroll 8
    for (int c = cb; c < ce; ++c)
      tot += min(__hip_atomic_load(&cell_count[k.c0 + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                 k.cell_cap);
  }
  int x = tot;  // inclusive scan of the thread sums within the wave
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  int run = carry_s + x - tot;
  for (int w = 0; w < wid; ++w) run += wsum[w];
  if (keep) {
#pragma unroll
    for (int u = 0; u < FILL_KEEP; ++u)
      if (cb + u < ce) {
        cell_start[k.c0 + cb + u] = run;
        run += kept[u];
      }
    return;
  }
#pragma unroll 8
  for (int c = cb; c < ce; ++c) {
    cell_start[k.c0 + c] = run;
    run += min(__hip_atomic_load(&cell_count[k.c0 + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
               k.cell_cap);
  }
}

// Spatial order.  At each rebuild the particles of a replica are ordered by
// cell (cells x fastest, each cell by ascending id): perm[s] is the particle
// at sorted position s, sidx[i] the sorted position of particle i, and
// [cell_start[c], cell_start[c] + cell_count[c]) the positions of cell c.
// The FP32 view, the list rows and the list entries use sorted positions, so
// a particle's neighbours lie close to it in memory; the FP64 state keeps the
// reference's particle order.
//
// Work items of a grid-strided rebuild kernel: workgroup b takes items b,
// b + gridDim.x, ... (at most 256 of them; the grid is sized so).  Which of
// them lie in a replica whose rebuild is live is found with one flag read
// per item, all in flight at once (thread t reads item b + t gridDim.x),
// so a step without a rebuild costs each workgroup one round trip.  An item
// is `cells` consecutive cells (it may straddle two replicas: both ends are
// read).  Returns the items' bits in live[4] (LDS), uniform after the barrier.
__device__ __forceinline__ void live_items(int n_items, int cells, int n_cells, int n_rep,
                                           const int* __restrict__ cell_rep,
                                           const int* __restrict__ flags_all,
                                           unsigned long long* live) {
  const int it = blockIdx.x + threadIdx.x * gridDim.x;
  bool on = false;
  if (it < n_items) {
    const int c0 = it * cells, c1 = min(c0 + cells, n_cells) - 1;
    const int ra = n_rep == 1 ? 0 : cell_rep[c0], rb = n_rep == 1 ? 0 : cell_rep[c1];
    on = (flags_all[ra * F_NFLAGS + F_LIVE] != 0 && flags_all[ra * F_NFLAGS + F_ERROR] == 0) ||
         (flags_all[rb * F_NFLAGS + F_LIVE] != 0 && flags_all[rb * F_NFLAGS + F_ERROR] == 0);
  }
  const unsigned long long bits = __builtin_amdgcn_ballot_w64(on);
  if ((threadIdx.x & 63) == 0) live[threadIdx.x >> 6] = bits;
  __syncthreads();
}

// Grid of a grid-strided rebuild kernel over n_items: at most `cap`
// workgroups, and at least enough that none takes more than 256 items.
static int strided_grid(int n_items, int cap) {
  const int need = (n_items + MFA_BLOCK - 1) / MFA_BLOCK;
  const int g = n_items < cap ? n_items : cap;
  return g > need ? g : need;
}

// One wave per cell, four cells per workgroup: orders each cell's particles
// by id, records the permutation, writes the FP32 view of each (relative to
// its cell's origin) at its sorted position and resets the skin reference
// (binning.f90:125-127).  A workgroup takes the groups of four cells
// blockIdx.x, blockIdx.x + gridDim.x, ..., so that a step without a rebuild
// costs a small launch even for a replica batch's many cells.
__device__ __forceinline__ void sort_cell(
    int cg, int n_cells, int lane, int* ids, const KParams* __restrict__ kb,
    const int* __restrict__ cell_rep, const double* __restrict__ r, const double* __restrict__ v,
    const int* __restrict__ type, const int* __restrict__ mic, const int* __restrict__ cell_count,
    const int* __restrict__ cell_slots, const int* __restrict__ cell_start, int* __restrict__ perm,
    int* __restrict__ sidx, double* __restrict__ r_ref, const int* __restrict__ flags_all,
    float4* __restrict__ pos4, float4* __restrict__ vel4) {
  const KParams& k = kb[cell_rep[cg < n_cells ? cg : 0]];
  const bool on = cg < n_cells && flags_all[k.rep * F_NFLAGS + F_LIVE] != 0 &&
                  flags_all[k.rep * F_NFLAGS + F_ERROR] == 0;
  const int m = on ? min(cell_count[cg], k.cell_cap) : 0;
  for (int t = lane; t < m; t += 64) ids[t] = cell_slots[(size_t)cg * k.cell_cap + t];
  __syncthreads();
  if (on) {
    const int c = cg - k.c0;
    const int s0 = cell_start[cg];
    for (int t = lane; t < m; t += 64) {
      const int id = ids[t];
      int rank = 0;
      for (int u = 0; u < m; ++u) rank += ids[u] < id ? 1 : 0;
      const int s = s0 + rank;
      perm[s] = id;
      sidx[id] = s;
      double ri[3], vi[3];
      for (int d = 0; d < 3; ++d) {
        ri[d] = r[3 * (size_t)id + d];
        vi[d] = v[3 * (size_t)id + d];
        r_ref[3 * (size_t)id + d] = ri[d] + (double)mic[3 * (size_t)id + d] * k.L[d];
      }
      write_view(k, s, c, type[id] - 1, id - k.p0, ri, vi, pos4, vel4);
    }
  }
  __syncthreads();  // the next group of cells reuses the ids
}

#define SORT_GRID 8192
__global__ __launch_bounds__(MFA_BLOCK) void k_cell_sort(
    const KParams* __restrict__ kb, const int* __restrict__ cell_rep, int n_cells, int n_rep,
    const double* __restrict__ r, const double* __restrict__ v, const int* __restrict__ type,
    const int* __restrict__ mic, const int* __restrict__ cell_count,
    const int* __restrict__ cell_slots, const int* __restrict__ cell_start,
    int* __restrict__ perm, int* __restrict__ sidx, double* __restrict__ r_ref,
    const int* __restrict__ flags_all, float4* __restrict__ pos4, float4* __restrict__ vel4) {
  extern __shared__ int sort_ids[];  // [4][cell_cap]
  __shared__ unsigned long long live[MFA_BLOCK / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int cell_cap = kb[0].cell_cap;  // the same for every replica
  const int n_grp = (n_cells + 3) / 4;
  if ((int)gridDim.x >= n_grp) {  // one group per workgroup: it checks its own replica
    sort_cell((int)blockIdx.x * 4 + w, n_cells, lane, sort_ids + w * cell_cap, kb, cell_rep, r, v,
              type, mic, cell_count, cell_slots, cell_start, perm, sidx, r_ref, flags_all, pos4, vel4);
    return;
  }
  live_items(n_grp, 4, n_cells, n_rep, cell_rep, flags_all, live);
  for (int q = 0; q < MFA_BLOCK / 64; ++q) {
    unsigned long long b = __builtin_amdgcn_readfirstlane(live[q] & 0xffffffffull) |
                           ((unsigned long long)__builtin_amdgcn_readfirstlane(live[q] >> 32) << 32);
    while (b) {
      const int t = q * 64 + __builtin_ctzll(b);
      b &= b - 1;
      const int grp = blockIdx.x + t * gridDim.x;
      sort_cell(grp * 4 + w, n_cells, lane, sort_ids + w * cell_cap, kb, cell_rep, r, v, type, mic,
                cell_count, cell_slots, cell_start, perm, sidx, r_ref, flags_all, pos4, vel4);
    }
  }
}

// Full Verlet list: every j != i in the 27 cells around i with
// r^2 < list_2(ti,tj) (binning.f90:64-113, both directions stored).
// One workgroup per cell, four waves; wave w takes the cell's particles in
// octants 2w and 2w+1 of the cell (two groups; a particle's row length lives
// in the lane of its place in the wave's list).  The candidates of the 27
// cells, in offset-code order 9(dx+1)+3(dy+1)+(dz+1) and, within a cell, in
// ascending id (the order of a sequential scan), are staged once per
// workgroup in LDS (position relative to this cell's origin, entry), in
// pieces of at most LB_PIECE; every wave then walks them 64 at a time.  A
// candidate farther than the list radius from the bounding box of a group's
// particles cannot be a neighbour of any of them; the others are queued (as
// indices into the staged piece), in order, for that group.  Each time 64
// are queued, and at the end of each piece, the group's particles are tested
// in turn against the queued candidates at once and a ballot appends the
// hits to the particle's row in lane order: the upper ones (neighbours at
// higher sorted positions) from the front, the lower ones from the back, so
// the upper part of every row holds its neighbours in scan order, ascending
// in entry value (code, then sorted position), which lets the half-list
// pass find a pair's place in the other row by bisection.  Separations come from the FP32 view (both
// positions relative to their cell origins, plus the offset of the
// candidate's cell): no periodic arithmetic.  A separation whose r^2 lies
// within the view's rounding bound of the radius is decided again in FP64
// with the reference's image convention (binning.f90:84-105).  Each entry is
// the candidate's sorted position with its cell's offset code above bit 27.
// With walls (no periodicity in z) cells beyond the z faces hold no
// neighbours.  A cell with fewer than LB_MIN_OCTANT particles, or with more
// than 32 in one octant, takes the same path with one group per wave
// (particles w + 4t) and no box test.
#define LB_WAVES 4                   // waves per cell (MFA_BLOCK / 64)
#define LB_MAX_CELL (64 * LB_WAVES)  // particles per cell the build takes
#define LB_RING 128                  // queued candidates per group (a block of 64 adds at most 64)
#define LB_MIN_OCTANT 16             // particles in a cell for the octant groups
#define LB_PIECE 1024                // candidates staged in LDS at a time
template <int KIND>
__device__ __forceinline__ void build_cell(
    int cg, const KParams* __restrict__ kb, const int* __restrict__ cell_rep,
    const double* __restrict__ r, const int* __restrict__ cell_count,
    const int* __restrict__ cell_start, const int* __restrict__ perm,
    const float4* __restrict__ pos4, unsigned* __restrict__ nl, int* __restrict__ nn, int cap,
    int* __restrict__ flags_all, int octants, int piece) {
  // dynamic LDS: the cell's own particles [cell_cap]; the staged candidates
  // (view relative to this cell, .w type bits) [piece] and their entries
  // [piece]; with octants, per wave and group the queued candidate indices
  extern __shared__ float4 lds_lb[];
  __shared__ int q_first[28], q_start[27];
  __shared__ float4 q_off[27];
  // per type pair: list radius^2 and interaction range^2 (near/shell rows)
  __shared__ float2 l2tab[MFA_NTYPE * MFA_NTYPE];
  __shared__ int hlist[LB_WAVES][64];  // the wave's particles, group 0 then group 1
  __shared__ int crowded_s;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const KParams& k = kb[cell_rep[cg]];
  int* flags = flags_all + k.rep * F_NFLAGS;
  if (flags[F_LIVE] == 0 || flags[F_ERROR] != 0) return;  // uniform over the workgroup
  const int c = cg - k.c0;
  const int cx = c % k.nb[0], cy = (c / k.nb[0]) % k.nb[1], cz = c / (k.nb[0] * k.nb[1]);
  const int m = min(cell_count[cg], k.cell_cap);
  const int hs0 = cell_start[cg];
  float4* hpos = lds_lb;
  float4* cpos = lds_lb + k.cell_cap;
  unsigned* cent = (unsigned*)(cpos + piece);
  int* qidx_w = (int*)(cent + piece) + w * 2 * LB_RING;  // [2][LB_RING] of this wave
  if (threadIdx.x < 27) {
    const int t = threadIdx.x;
    const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
    const int jx = cx + dx, jy = cy + dy, jz = cz + dz;
    const int cc = k.c0 + (((jz + k.nb[2]) % k.nb[2]) * k.nb[1] + (jy + k.nb[1]) % k.nb[1]) * k.nb[0] +
                   (jx + k.nb[0]) % k.nb[0];
    const bool gone = k.symmetry == 0 && (jz < 0 || jz >= k.nb[2]);
    const int cnt = gone ? 0 : min(cell_count[cc], k.cell_cap);
    // inclusive prefix over the 27 codes (lanes 0..26 of wave 0)
    int x = cnt;
    for (int o = 1; o < 32; o <<= 1) {
      const int y = __shfl_up(x, o, 32);
      if (t >= o) x += y;
    }
    q_first[t + 1] = x;
    if (t == 0) q_first[0] = 0;
    q_start[t] = gone ? 0 : cell_start[cc];
    q_off[t] = make_float4(-(float)dx * (float)k.r_bin[0], -(float)dy * (float)k.r_bin[1],
                           -(float)dz * (float)k.r_bin[2], 0.0f);
  }
  if (threadIdx.x >= 32 && threadIdx.x < 32 + MFA_NTYPE * MFA_NTYPE) {
    const int q = threadIdx.x - 32;
    l2tab[q] = make_float2((float)k.list_2[q / MFA_NTYPE][q % MFA_NTYPE],
                           (float)(k.list_2[q / MFA_NTYPE][q % MFA_NTYPE] - k.skin_2));
  }
  if (threadIdx.x == 64) crowded_s = 0;
  for (int t = threadIdx.x; t < m; t += MFA_BLOCK) hpos[t] = pos4[hs0 + t];
  __syncthreads();
  if (c == 0 && threadIdx.x == 0) {
    atomicAdd(&flags[F_NREBUILD], 1);
    if (KIND == LK_HALF) atomicExch(&flags[F_MIRROR], 1);  // the half-list pass must find new mirror slots
  }
  // the wave's two octants: particle lists and their counts
  const float hx = 0.5f * (float)k.r_bin[0], hy = 0.5f * (float)k.r_bin[1],
              hz = 0.5f * (float)k.r_bin[2];
  int n0 = 0, n1 = 0;  // wave-uniform
  // sparse cells (a few particles per wave) gain nothing from the box test
  const bool sparse = !octants || m < LB_MIN_OCTANT;
  for (int h0 = 0; h0 < (sparse ? 0 : m); h0 += 64) {
    const int h = h0 + lane;
    int oc = -1;
    if (h < m) {
      const float4 p = hpos[h];
      oc = (p.x >= hx ? 1 : 0) | (p.y >= hy ? 2 : 0) | (p.z >= hz ? 4 : 0);
    }
    const unsigned long long m0 = __builtin_amdgcn_ballot_w64(oc == 2 * w);
    const unsigned long long m1 = __builtin_amdgcn_ballot_w64(oc == 2 * w + 1);
    if (oc == 2 * w || oc == 2 * w + 1) {
      const unsigned long long mm = oc == 2 * w ? m0 : m1;
      const int at = (oc == 2 * w ? n0 : 32 + n1) +
                     (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mm, 0u));
      if (at < 64 && (oc == 2 * w ? at < 32 : true)) hlist[w][at] = h;
    }
    n0 += (int)__popcll(m0);
    n1 += (int)__popcll(m1);
  }
  if (lane == 0 && (n0 > 32 || n1 > 32)) atomicOr(&crowded_s, 1);
  __syncthreads();
  const bool crowded = sparse || crowded_s != 0;
  // groups: [first particle in hlist, count]; a crowded cell takes particles
  // w + 4t as one group with no box test
  int g_n[2], g_at[2];
  if (crowded) {
    g_n[0] = (m - w + LB_WAVES - 1) / LB_WAVES;
    g_n[1] = 0;
    g_at[0] = 0;
    g_at[1] = 0;
  } else {
    g_n[0] = n0;
    g_n[1] = n1;
    g_at[0] = 0;
    g_at[1] = 32;
  }
  // a wave without particles still stages candidates and meets the barriers
  const bool busy = g_n[0] + g_n[1] > 0;
  const int total = q_first[27];
  const float margin = k.lf_margin;
  float l2max = 0.0f;
  for (int q = 0; q < MFA_NTYPE * MFA_NTYPE; ++q) l2max = fmaxf(l2max, l2tab[q].x);
  const float reach2 = l2max + 2.0f * margin;
  // each group's box: the bounding box of its particles' view positions
  float bmin[2][3], bmax[2][3];
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    if (crowded) break;
    const bool mine = lane < g_n[g];
    float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (mine) p = hpos[hlist[w][g_at[g] + lane]];
    float lo[3] = {mine ? p.x : 3.0e38f, mine ? p.y : 3.0e38f, mine ? p.z : 3.0e38f};
    float hi[3] = {mine ? p.x : -3.0e38f, mine ? p.y : -3.0e38f, mine ? p.z : -3.0e38f};
    for (int o = 32; o > 0; o >>= 1)
      for (int d = 0; d < 3; ++d) {
        lo[d] = fminf(lo[d], __shfl_xor(lo[d], o, 64));
        hi[d] = fmaxf(hi[d], __shfl_xor(hi[d], o, 64));
      }
    for (int d = 0; d < 3; ++d) {
      bmin[g][d] = lo[d];
      bmax[g][d] = hi[d];
    }
  }
  int hc = 0;  // row length of the wave's particle in this lane's place
  int hu = 0;  // ... and how many of its entries are upper ones

  // Tests 64 candidates (one per lane; view position relative to this cell,
  // entry) against each particle of group g and appends the hits.
  auto test_block = [&](int g, const float4 cp, const unsigned entry, const bool live) {
    const int js_c = (int)(entry & MFA_JMASK);
    const float far = live ? 0.0f : 1.0e30f;
    const float pjx = cp.x + far, pjy = cp.y, pjz = cp.z;
    const int tj = __float_as_int(cp.w) & 3;
    for (int t = 0; t < g_n[g]; ++t) {
      const int h = __builtin_amdgcn_readfirstlane(crowded ? w + LB_WAVES * t : hlist[w][g_at[g] + t]);
      const int slot = (g == 0 ? 0 : g_n[0]) + t;
      const float4 pi = hpos[h];
      const int ti = __float_as_int(pi.w) & 3;
      const float dx = pi.x - pjx, dy = pi.y - pjy, dz = pi.z - pjz;
      const float d2 = dx * dx + dy * dy + dz * dz;
      const float2 l2n2 = l2tab[ti * MFA_NTYPE + tj];
      const float l2 = l2n2.x;
      const bool cand = live && js_c != hs0 + h;
      bool hit = cand && d2 < l2 - margin;
      const bool band = cand && !(d2 < l2 - margin) && d2 < l2 + margin;
      // rare: a separation within the view's rounding bound of the radius is
      // decided in FP64 (one wave-uniform test keeps the common path free of
      // nested divergence)
      if (__builtin_expect(__builtin_amdgcn_ballot_w64(band) != 0ull, 0)) {
        if (band) {
          const int i = perm[hs0 + h], j = perm[js_c];
          double ex = r[3 * (size_t)i + 0] - r[3 * (size_t)j + 0];
          double ey = r[3 * (size_t)i + 1] - r[3 * (size_t)j + 1];
          double ez = r[3 * (size_t)i + 2] - r[3 * (size_t)j + 2];
          min_image(k, ex, ey, ez);
          const double r2 = (ex * ex + ey * ey) + ez * ez;
          hit = r2 < k.list_2[ti][tj];
        }
      }
      const unsigned long long mask = __builtin_amdgcn_ballot_w64(hit);
      if (mask) {
        const int base = __builtin_amdgcn_readlane(hc, slot);
        if constexpr (KIND != LK_PLAIN) {
          // split rows: half-list rows put the upper entries (a neighbour at
          // a higher sorted position) in the front, the lower ones in the
          // back; near/shell rows put the entries inside the interaction
          // range in the front, those in the skin in the back (an FP32 test:
          // the split only steers how the force pass walks the row); both
          // parts in lane order
          const bool up = KIND == LK_HALF ? js_c > hs0 + h : d2 < l2n2.y;
          const unsigned long long mu = __builtin_amdgcn_ballot_w64(hit && up);
          const int bup = __builtin_amdgcn_readlane(hu, slot);
          const int tot = base + (int)__popcll(mask);
          if (hit && tot <= cap) {  // beyond the capacity: an overflow, reported below
            const unsigned long long mm = up ? mu : mask & ~mu;
            const int rank = (int)__builtin_amdgcn_mbcnt_hi(
                (unsigned)(mm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mm, 0u));
            const int pos = up ? bup + rank : cap - 1 - ((base - bup) + rank);
            nl[(size_t)(hs0 + h) * cap + pos] = entry;
          }
          hu = lane == slot ? bup + (int)__popcll(mu) : hu;
        } else {
          if (hit) {
            const unsigned pos = __builtin_amdgcn_mbcnt_hi(
                (unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, (unsigned)base));
            if (pos < (unsigned)cap) nl[(size_t)(hs0 + h) * cap + pos] = entry;
          }
        }
        hc = lane == slot ? base + (int)__popcll(mask) : hc;
      }
    }
  };
  // The queued candidates [hd, hd + cnt) of group g (cnt <= 64): indices
  // into the staged piece (a lane past cnt reads index 0 and is not live).
  auto drain = [&](int g, int hd, int cnt) {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int e = lane < cnt ? qidx_w[g * LB_RING + ((hd + lane) & (LB_RING - 1))] : 0;
    test_block(g, cpos[e], cent[e], lane < cnt);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  };

  for (int p0 = 0; p0 < total; p0 += piece) {  // uniform over the workgroup
    const int np = min(piece, total - p0);
    // stage candidates [p0, p0 + np): code q of candidate e is the last q with
    // q_first[q] <= e (binary search), its sorted position follows
    for (int t = threadIdx.x; t < np; t += MFA_BLOCK) {
      const int e = p0 + t;
      int lo = 0, hi = 27;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (q_first[mid] <= e) lo = mid;
        else hi = mid;
      }
      const int js = q_start[lo] + (e - q_first[lo]);
      const float4 pj = pos4[js];
      const float4 o = q_off[lo];
      cpos[t] = make_float4(pj.x - o.x, pj.y - o.y, pj.z - o.z, pj.w);
      cent[t] = (unsigned)js | ((unsigned)lo << MFA_JBITS);
    }
    __syncthreads();
    if (busy) {
      int head[2] = {0, 0}, tail[2] = {0, 0};  // wave-uniform queue positions
      for (int e0 = 0; e0 < np; e0 += 64) {
        const int e = e0 + lane;
        const bool live = e < np;
        const int ec = live ? e : 0;
        const float4 cp = cpos[ec];
        if (crowded) {  // one group, every candidate
          test_block(0, cp, cent[ec], live);
          continue;
        }
#pragma unroll
        for (int g = 0; g < 2; ++g) {
          if (g_n[g] == 0) continue;
          const float bx = fmaxf(fmaxf(bmin[g][0] - cp.x, cp.x - bmax[g][0]), 0.0f);
          const float by = fmaxf(fmaxf(bmin[g][1] - cp.y, cp.y - bmax[g][1]), 0.0f);
          const float bz = fmaxf(fmaxf(bmin[g][2] - cp.z, cp.z - bmax[g][2]), 0.0f);
          const bool pass = live && bx * bx + by * by + bz * bz < reach2;
          const unsigned long long pm = __builtin_amdgcn_ballot_w64(pass);
          if (pass) {
            const int at = (tail[g] + (int)__builtin_amdgcn_mbcnt_hi(
                                          (unsigned)(pm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)pm, 0u))) &
                           (LB_RING - 1);
            qidx_w[g * LB_RING + at] = e;
          }
          tail[g] += (int)__popcll(pm);
          if (tail[g] - head[g] >= 64) {
            drain(g, head[g], 64);
            head[g] += 64;
          }
        }
      }
      // the queues refer to this piece: test what is left before it goes
#pragma unroll
      for (int g = 0; g < 2; ++g)
        if (tail[g] > head[g]) drain(g, head[g], tail[g] - head[g]);
    }
    __syncthreads();  // every wave is done with the piece
  }
  if (!busy) return;
  const int nw = g_n[0] + g_n[1];
  int my_max = 0;
  if (lane < nw) {
    const int g = lane < g_n[0] ? 0 : 1, t = lane - (g == 0 ? 0 : g_n[0]);
    const int h = crowded ? w + LB_WAVES * t : hlist[w][g_at[g] + t];
    // unsplit rows: every entry counts as upper (list_slot(q) = q)
    nn[hs0 + h] = hc <= cap ? (hc | ((KIND != LK_PLAIN ? hu : hc) << 16)) : cap;
    if (hc > cap) atomicCAS(&flags[F_ERROR], 0, MFA_ERR_LIST_OVERFLOW);
    my_max = hc;
  }
  for (int o = 32; o > 0; o >>= 1) my_max = max(my_max, __shfl_xor(my_max, o, 64));
  // the maximum only grows: most builds find it already at least this large
  // and take no atomic on the one shared address
  if (lane == 0 && my_max > __hip_atomic_load(&flags[F_MAXNEIGH], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(&flags[F_MAXNEIGH], my_max);
}

// The list build's launch: a grid of at most LB_GRID workgroups (more if a
// workgroup would take over 256 cells), each taking the cells blockIdx.x,
// blockIdx.x + gridDim.x, ... of replicas with a live rebuild in turn, so
// that a step without a rebuild costs a small launch even for the hundreds
// of thousands of cells of a replica batch.
#define LB_GRID 8192
template <int KIND>
__global__ __launch_bounds__(MFA_BLOCK) void k_build_list(
    int n_cells, int n_rep, const KParams* __restrict__ kb, const int* __restrict__ cell_rep,
    const double* __restrict__ r, const int* __restrict__ cell_count,
    const int* __restrict__ cell_start, const int* __restrict__ perm,
    const float4* __restrict__ pos4, unsigned* __restrict__ nl, int* __restrict__ nn, int cap,
    int* __restrict__ flags_all, int octants, int piece) {
  __shared__ unsigned long long live[MFA_BLOCK / 64];
  if ((int)gridDim.x >= n_cells) {  // one cell per workgroup: it checks its own replica
    build_cell<KIND>(blockIdx.x, kb, cell_rep, r, cell_count, cell_start, perm, pos4, nl, nn, cap,
                      flags_all, octants, piece);
    return;
  }
  live_items(n_cells, 1, n_cells, n_rep, cell_rep, flags_all, live);
  for (int q = 0; q < MFA_BLOCK / 64; ++q) {
    unsigned long long b = __builtin_amdgcn_readfirstlane(live[q] & 0xffffffffull) |
                           ((unsigned long long)__builtin_amdgcn_readfirstlane(live[q] >> 32) << 32);
    while (b) {
      const int t = q * 64 + __builtin_ctzll(b);
      b &= b - 1;
      build_cell<KIND>(blockIdx.x + t * gridDim.x, kb, cell_rep, r, cell_count, cell_start, perm,
                        pos4, nl, nn, cap, flags_all, octants, piece);
      __syncthreads();  // the next cell reuses the LDS
    }
  }
}

// ---------------------------------------------------------------------------
// Force pass.  A lane group of TPP lanes handles one particle i: the lanes
// split i's neighbour list, then combine their sums with a fixed butterfly,
// so the result is independent of timing.  Everything acting on i is summed
// here (pairs, walls, bonds), so no atomics are needed and F is written once.
//
// Plain FP64 (k_force64): the reference's arithmetic for LJ
// (fluid_fluid.f90:94-126), DPD (dpd_forces.f90:19-94), 9-3 walls
// (fluid_wall.f90:81-121) and FENE (intra_molec.f90:16-139), summed in FP64.
//
// Mixed (k_force_mixed): separations from the FP32 cell-relative view plus
// the FP32 offset of the neighbour's cell (no periodic arithmetic per pair);
// a separation whose r^2 lies within the view's error bound of a cut-off is
// decided again in FP64 with the reference's image convention, so the
// interacting set is the FP64 one.  Pair math in FP32; each pair term enters
// an exact fixed-point sum.

// Adds a pair/bond/wall term to the fixed-point sums.  Each component is
// truncated toward zero to a multiple of 2^-22 and added as an
// integer-valued double: such sums are exact below 2^53, so they do not
// depend on the order of the terms (they equal the int64 sums, with one
// FP64 add per component).  A lane sum at or beyond 2^52 (or not finite) is
// reported by the caller as a bad force.
__device__ __forceinline__ void fix_add3(double& ax, double& ay, double& az, float fx, float fy,
                                         float fz) {
  ax += (double)truncf(fx * MFA_FIX_SCALE);
  ay += (double)truncf(fy * MFA_FIX_SCALE);
  az += (double)truncf(fz * MFA_FIX_SCALE);
}
__device__ __forceinline__ bool fix_bad(double ax, double ay, double az) {
  const double lim = 4503599627370496.0;  // 2^52
  return !(fabs(ax) < lim && fabs(ay) < lim && fabs(az) < lim);
}

template <int TPP>
__device__ __forceinline__ double group_sum(double v) {
  for (int o = TPP / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <int TPP>
__device__ __forceinline__ long long group_sum(long long v) {
  for (int o = TPP / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Normal deviate of pair (i,j): FP64 and FP32 versions of the same transform.
__device__ __forceinline__ double pair_normal_f64(uint32_t seed, int64_t step, int i, int j) {
  uint32_t w0, w1;
  mfa_pair_words(seed, step, i, j, &w0, &w1);
  const double u1 = (double)(w0 >> 8) * 0x1p-24 + 0x1p-25;
  const double u2 = (double)(w1 >> 8) * 0x1p-24;
  // cos(2 pi u2) as cospi(2 u2): the argument is reduced exactly, within an
  // ulp or two of cos(fl(2 pi) u2)
  return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}
__device__ __forceinline__ float pair_normal_f32(uint32_t seed, int64_t step, int i, int j) {
  uint32_t w0, w1;
  mfa_pair_words(seed, step, i, j, &w0, &w1);
  const float u1 = (float)(w0 >> 8) * 0x1p-24f + 0x1p-25f;
  const float u2 = (float)(w1 >> 8) * 0x1p-24f;
  // v_log_f32 is log2; v_cos_f32 takes revolutions (cos(2 pi x)).
  const float l2 = __builtin_amdgcn_logf(u1);
  return __builtin_amdgcn_sqrtf(-1.3862943611198906f * l2) * __builtin_amdgcn_cosf(u2);
}

// Langevin force on particle il (THERMOSTAT 1; lgv_forces.f90:25-43, called
// for every particle at the end of its fluid_fluid row, fluid_fluid.f90:219-221):
// per component (sig * xi_d) * sqrt(m) - (gamma_1 m) v_d, xi_d the particle's
// counter deviate d (the reference draws rnor() three times).
__device__ __forceinline__ void langevin_force(const KParams& k, int64_t step, int il, int ti,
                                               double vx, double vy, double vz, double* fl) {
  const double m = k.mass_type[ti], sm = sqrt(m);
  const double vv[3] = {vx, vy, vz};
  for (int d = 0; d < 3; ++d) {
    uint32_t w0, w1;
    mfa_particle_words(k.seed, step, il, d, &w0, &w1);
    const double u1 = (double)(w0 >> 8) * 0x1p-24 + 0x1p-25;
    const double u2 = (double)(w1 >> 8) * 0x1p-24;
    const double xi = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    fl[d] = (k.sig * xi) * sm - (k.friction1 * m) * vv[d];
  }
}

// Switch-on clamp of r^2 (fluid_fluid.f90:42-49) with r_time = step*dt
// (md_main.f90:69, step = s_time + i_time - 1); 0 once the ramp is over.
__device__ __forceinline__ double switch_on_r2min(const KParams& k, int64_t step) {
  if (k.fso_time > 0.0) {
    const double r_time = (double)step * k.dt;
    if (r_time < k.fso_time) return (1.0 - r_time / k.fso_time) * k.fso_r2_init;
  }
  return 0.0;
}

// The chain of particle i and its length: brush [0,n_brush) with n_mon, melt
// [part_init_d,part_init_e) with n_mon_d, particle-4 chains [part_init_e,n)
// with n_mon_e (intra_molec.f90:16-139).
__device__ __forceinline__ void chain_of(const KParams& k, int i, int& base, int& len) {
  base = -1;
  len = 0;
  if (i < k.n_brush) {
    base = (i / k.n_mon) * k.n_mon;
    len = k.n_mon;
  } else if (i < k.part_init_e) {
    base = k.part_init_d + ((i - k.part_init_d) / k.n_mon_d) * k.n_mon_d;
    len = k.n_mon_d;
  } else if (k.n_mon_e > 0) {
    base = k.part_init_e + ((i - k.part_init_e) / k.n_mon_e) * k.n_mon_e;
    len = k.n_mon_e;
  }
}

// 9-3 walls at z = 0 and z = z_space_wall (fluid_wall.f90:81-121); the energy
// uses a_w, the force a_wall(type) (:98,101,116,121).
__device__ __forceinline__ void wall_93(const KParams& k, int ti, double zi, double& fb,
                                        double& ft, double& e_fw) {
  const double sw = k.sigma_wall[ti], aw = k.a_wall[ti], aabs = fabs(aw);
  const double sw3 = pow3(sw), sw9 = pow3(sw3);
  double inv_z = 1.0 / zi;
  double rd = sw * inv_z;
  double rd3 = pow3(rd);
  e_fw += aabs * pow3(rd3) - k.a_w * rd3;
  double iz2 = inv_z * inv_z, iz4 = iz2 * iz2, iz8 = iz4 * iz4;
  fb = (9.0 * aabs) * sw9 * (iz8 * iz2) - (3.0 * aw) * sw3 * iz4;
  inv_z = 1.0 / (k.z_space_wall - zi);
  rd = sw * inv_z;
  rd3 = pow3(rd);
  e_fw += aabs * pow3(rd3) - k.a_w * rd3;
  iz2 = inv_z * inv_z;
  iz4 = iz2 * iz2;
  iz8 = iz4 * iz4;
  ft = -(9.0 * aabs) * sw9 * (iz8 * iz2) + (3.0 * aw) * sw3 * iz4;
}

// Minimum image in x, y and z as bending.f90:48-52 writes it (also with walls).
__device__ __forceinline__ void min_image3(const KParams& k, double* d) {
  for (int j = 0; j < 3; ++j) d[j] = pbc(d[j], k.L[j], k.invL[j]);
}

// The angle of a brush chain centred at bead `cen` (bending.f90:43-118):
// fp = F_mod dir_prev (on cen-1), fn = F_mod dir_next (on cen+1); cen itself
// takes -(fp + fn).  Returns 0.5 F_mod delta_alpha.
__device__ double bend_angle(const KParams& k, const double* __restrict__ r, int cen, double* fp,
                             double* fn) {
  double rp[3], rn[3];
  for (int j = 0; j < 3; ++j) {
    rp[j] = r[3 * (size_t)cen + j] - r[3 * (size_t)(cen - 1) + j];
    rn[j] = r[3 * (size_t)(cen + 1) + j] - r[3 * (size_t)cen + j];
  }
  min_image3(k, rp);
  min_image3(k, rn);
  double dot_pr = 0.0, rp2 = 0.0, rn2 = 0.0, dp2 = 0.0, dn2 = 0.0;
  for (int j = 0; j < 3; ++j) {
    dot_pr = dot_pr + rn[j] * rp[j];
    rp2 = rp2 + rp[j] * rp[j];
    rn2 = rn2 + rn[j] * rn[j];
  }
  const double cos_alpha = dot_pr / sqrt(rp2 * rn2);
  double dn[3], dp[3];
  for (int j = 0; j < 3; ++j) {
    dn[j] = rp[j] * rn2 - rn[j] * dot_pr;
    dp[j] = -rn[j] * rp2 + rp[j] * dot_pr;
  }
  for (int j = 0; j < 3; ++j) {
    dp2 = dp2 + dp[j] * dp[j];
    dn2 = dn2 + dn[j] * dn[j];
  }
  const double eps = (double)0.0000001f;  // a default-real literal (bending.f90:72,81)
  if (dn2 < eps) {
    for (int j = 0; j < 3; ++j) dn[j] = dn[j] * 0.0;
  } else {
    const double nn = sqrt(dn2 * rn2);
    for (int j = 0; j < 3; ++j) dn[j] = dn[j] / nn;
  }
  if (dp2 < eps) {
    for (int j = 0; j < 3; ++j) dp[j] = dp[j] * 0.0;
  } else {
    const double np_ = sqrt(dp2 * rp2);
    for (int j = 0; j < 3; ++j) dp[j] = dp[j] / np_;
  }
  const double delta_alpha = cos_alpha >= 1.0 ? 0.0 : acos(cos_alpha) - k.alpha_eq;
  const double f_mod = k.k_bend * delta_alpha;
  for (int j = 0; j < 3; ++j) {
    fp[j] = f_mod * dp[j];
    fn[j] = f_mod * dn[j];
  }
  return .5 * f_mod * delta_alpha;
}

// What the force pass of a step that also integrates (either form) reads and writes
// besides r, the view and the list: the next positions and velocities (the
// other buffers: the FP64 pass reads its neighbours' velocities), wrap counts
// (in the step: in place, mic_out = mic_in), the next FP32 view (the other
// view buffer), the skin test's rebuild flags (one F_NFLAGS row per replica)
// and zeroed cell counts.
struct FuseArgs {
  double* r_out;
  const double* v_in;
  double* v_out;
  const int* mic_in;
  int* mic_out;
  const double* r_ref;
  const int* cell_of;
  float4* pos4_out;
  float4* vel4_out;
  int* cell_count;
  int* rebuild_flags;
  int nc_tot;
  // charged systems: the k-space force partials of this step's positions
  // (k_ewald_force*) added to F before the integration, as k_ewald_apply
  // adds them to F otherwise; qinv[i] = charged index of particle i or -1
  // (qinv null: no k-space term)
  const int* qinv;
  const double* qval;
  const double* fk_part;
  const double* mz;
  double inv_V;
  int nq, n_kgroups, add_dip;
};

// The half-list form's buffers: per list slot, the scaled pair scalar of an
// upper entry (written by the first pass, 0 where the pair does not
// interact) and, for a lower entry, the slot of its pair in the neighbour's
// row; per particle (sorted order), the first pass's fixed-point sums.
struct HalfArgs {
  float* pair_s;
  uint16_t* mir;
  double* fpart;
};

// The k-space force partials: per charged particle idx, per force group kg
// (a block row of the force kernel), three components, particle-major, so
// that one particle's partials lie together.
__device__ __forceinline__ size_t fk_index(int idx, int kg, int n_kgroups) {
  return ((size_t)idx * (size_t)n_kgroups + (size_t)kg) * 3;
}

// Sum of the partials of charged particle idx, component d, in a fixed
// order that a lane group can share: eight strands (strand l: groups l,
// l + 8, l + 16, ... added in turn), then ((t0 + t4) + (t2 + t6)) +
// ((t1 + t5) + (t3 + t7)).
__device__ __forceinline__ double fk_strand(const double* __restrict__ fk_part, int idx, int n_kgroups,
                                            int d, int l) {
  const double* p = fk_part + fk_index(idx, 0, n_kgroups) + d;
  double t = 0.0;
  for (int kg = l; kg < n_kgroups; kg += 8) t += p[3 * (size_t)kg];
  return t;
}
__device__ __forceinline__ double fk_sum(const double* __restrict__ fk_part, int idx, int n_kgroups,
                                         int d) {
#pragma clang fp contract(off)
  double t[8];
  for (int l = 0; l < 8; ++l) t[l] = fk_strand(fk_part, idx, n_kgroups, d, l);
  return ((t[0] + t[4]) + (t[2] + t[6])) + ((t[1] + t[5]) + (t[3] + t[7]));
}
// The same sum by the TPP lanes of a group (every lane gets it; idx < 0:
// a particle without charge, 0).  Lane sub holds strands sub, sub + TPP, ...
// (mod 8), pairs them as the fixed order does and the butterfly finishes it.
template <int TPP>
__device__ __forceinline__ double fk_sum_group(const double* __restrict__ fk_part, int idx,
                                               int n_kgroups, int d, int sub) {
#pragma clang fp contract(off)
  auto st = [&](int l) { return idx >= 0 ? fk_strand(fk_part, idx, n_kgroups, d, l) : 0.0; };
  if constexpr (TPP == 1) {
    return idx >= 0 ? fk_sum(fk_part, idx, n_kgroups, d) : 0.0;
  } else if constexpr (TPP == 2) {
    const double b = (st(sub) + st(sub + 4)) + (st(sub + 2) + st(sub + 6));
    return b + __shfl_xor(b, 1, 64);
  } else if constexpr (TPP == 4) {
    const double a = st(sub) + st(sub + 4);
    const double b = a + __shfl_xor(a, 2, 64);
    return b + __shfl_xor(b, 1, 64);
  } else {
    const double t = st(sub & 7);
    const double a = t + __shfl_xor(t, 4, 64);
    const double b = a + __shfl_xor(a, 2, 64);
    return b + __shfl_xor(b, 1, 64);
  }
}

// F_i,d plus the k-space force on a charged particle from its partial sum s:
// -8 pi q s / V (ewald_k.f90:177-192), and with add_dip the slab dipole force
// -4 pi q M_z / V on z (dipolar_correction.f90:26-30).  The same arithmetic,
// uncontracted, wherever it is added.
__device__ __forceinline__ double kspace_add(double fd, double s, double q, double inv_V, int add_k,
                                             int add_dip, int d, const double* __restrict__ mz) {
#pragma clang fp contract(off)
  const double pi_f = (double)3.1415926536f;  // mfa_common.f90:332
  if (add_k) fd += -8.0 * pi_f * q * s * inv_V;
  if (add_dip && d == 2) fd -= 4.0 * pi_f * q * (*mz) * inv_V;
  return fd;
}

// The end of a force pass that also integrates: verlet_velocities of this
// step with F_i (k_fused_integrate's order: a = F/m, v += dt/2 a, pinned
// heads), then verlet_positions and check_skin of the next.  Positions are
// read from r (this step's buffer, which the whole pass reads) and written
// to fu.r_out; only the group's lane 0 reads or writes particle i's v and
// mic.  Every lane of the launch takes part (ballot, count clearing).
// CHARGED: compiled with the k-space force's addition (kernels of systems
// with charges or bending; fu.qinv says whether there are charges).
template <int TPP, bool CHARGED, bool SPLIT>
__device__ __forceinline__ void fused_epilogue(const KParams& k, bool active, int sub, int i, int il,
                                               int ti, int ss, double fx, double fy, double fz,
                                               const double* __restrict__ r, const FuseArgs& fu) {
  bool over = false;
  if constexpr (CHARGED) {
    // the k-space force of this step, summed by the lane group; every lane
    // of the group loads the particle's state with its partials, so that the
    // two wait for memory once
    double vi[3], ri[3], rr[3];
    int mi[3];
    int qidx = -1;
    double fk[3] = {0.0, 0.0, 0.0};
    if (fu.qinv) {  // uniform over the launch
      qidx = active ? fu.qinv[i] : -1;
      if (active)
        for (int d = 0; d < 3; ++d) {
          vi[d] = fu.v_in[3 * (size_t)i + d];
          ri[d] = r[3 * (size_t)i + d];
          mi[d] = fu.mic_in[3 * (size_t)i + d];
          rr[d] = fu.r_ref[3 * (size_t)i + d];
        }
      for (int d = 0; d < 3; ++d) fk[d] = fk_sum_group<TPP>(fu.fk_part, qidx, fu.n_kgroups, d, sub);
    }
    if (active && sub == 0) {
      double fi[3] = {fx, fy, fz}, ai[3];
      if (!fu.qinv)
        for (int d = 0; d < 3; ++d) {
          vi[d] = fu.v_in[3 * (size_t)i + d];
          ri[d] = r[3 * (size_t)i + d];
          mi[d] = fu.mic_in[3 * (size_t)i + d];
          rr[d] = fu.r_ref[3 * (size_t)i + d];
        }
      if (qidx >= 0)
        for (int d = 0; d < 3; ++d)
          fi[d] = kspace_add(fi[d], fk[d], fu.qval[qidx], fu.inv_V, 1, fu.add_dip, d, fu.mz);
      vv_one(k, il, ti, fi, vi, ai);
      bool wrapped = false;
      vp_one(k, ri, vi, ai, mi, wrapped);
      for (int d = 0; d < 3; ++d) {
        fu.r_out[3 * (size_t)i + d] = ri[d];
        fu.v_out[3 * (size_t)i + d] = vi[d];
      }
      if (wrapped)
        for (int d = 0; d < 3; ++d) fu.mic_out[3 * (size_t)i + d] = mi[d];
      write_view(k, ss, fu.cell_of[i], ti, il, ri, vi, fu.pos4_out, fu.vel4_out);
      over = skin_one(k, ri, rr, mi);
    }
  } else if constexpr (TPP >= 4 && SPLIT) {
    // lanes 0, 1, 2 of the group integrate x, y, z (vv_one, vp_one and
    // write_view one component at a time, the same operations); lane 0
    // gathers the skin test's three terms and adds them in skin_one's order.
    // (The mixed pass only: in the FP64 pass the extra registers cost more
    // than the shorter tail saves.)
    double dd = 0.0;
    if (active && sub < 3) {
      const int d = sub;
      const size_t o = 3 * (size_t)i + d;
      const double Ld = d == 0 ? k.L[0] : (d == 1 ? k.L[1] : k.L[2]);
      double vd = fu.v_in[o], rd = r[o];
      int md = fu.mic_in[o];
      const double rrd = fu.r_ref[o];
      const int cell = fu.cell_of[i];
      const double fd = d == 0 ? fx : (d == 1 ? fy : fz);
      double ad = fd * k.inv_mass_type[ti];
      vd = vd + k.half_dt * ad;
      if (is_head(k, il) && !k.dpd_vv) {
        const int c = il / k.n_mon;
        const double sgn = c < k.n_chain / 2 ? 1.0 : -1.0;
        vd = d == 2 ? 0.0 : sgn * (d == 0 ? k.va[0] : k.va[1]);
        ad = 0.0;
      }
      rd = (rd + k.dt * vd) + k.half_dt_2 * ad;
      vd = vd + k.half_dt * ad;
      const bool wraps = d < 2 || k.symmetry != 0;
      bool wrapped = false;
      if (wraps) {
        if (rd > Ld) {
          rd = rd - Ld;
          md += 1;
          wrapped = true;
        } else if (rd < 0.0) {
          rd = rd + Ld;
          md -= 1;
          wrapped = true;
        }
      }
      fu.r_out[o] = rd;
      fu.v_out[o] = vd;
      if (wrapped) fu.mic_out[o] = md;
      const int cx = cell % k.nb[0], cy = (cell / k.nb[0]) % k.nb[1], cz = cell / (k.nb[0] * k.nb[1]);
      const int cd = d == 0 ? cx : (d == 1 ? cy : cz);
      const double bd = d == 0 ? k.r_bin[0] : (d == 1 ? k.r_bin[1] : k.r_bin[2]);
      const double iLd = d == 0 ? k.invL[0] : (d == 1 ? k.invL[1] : k.invL[2]);
      double ld = rd - (double)cd * bd;
      if (d < 2 || k.symmetry == 1) ld -= Ld * rint(ld * iLd);
      float* pv = reinterpret_cast<float*>(fu.pos4_out + ss);
      float* vv = reinterpret_cast<float*>(fu.vel4_out + ss);
      pv[d] = (float)ld;
      vv[d] = (float)vd;
      if (d == 0) {
        pv[3] = __int_as_float((il << 2) | ti);
        vv[3] = 0.0f;
      }
      dd = (rrd - rd) - (double)md * Ld;
    }
    const int base = (threadIdx.x & 63) & ~(TPP - 1);
    const double d1 = __shfl(dd, base + 1, 64), d2 = __shfl(dd, base + 2, 64);
    if (active && sub == 0) {
      double s = 0.0;
      s = s + dd * dd;
      s = s + d1 * d1;
      s = s + d2 * d2;
      over = s > k.skin_half_2;
    }
  } else {
    if (active && sub == 0) {
      double fi[3] = {fx, fy, fz}, vi[3], ai[3], ri[3], rr[3];
      int mi[3];
      for (int d = 0; d < 3; ++d) {
        vi[d] = fu.v_in[3 * (size_t)i + d];
        ri[d] = r[3 * (size_t)i + d];
        mi[d] = fu.mic_in[3 * (size_t)i + d];
        rr[d] = fu.r_ref[3 * (size_t)i + d];
      }
      vv_one(k, il, ti, fi, vi, ai);
      bool wrapped = false;
      vp_one(k, ri, vi, ai, mi, wrapped);
      for (int d = 0; d < 3; ++d) {
        fu.r_out[3 * (size_t)i + d] = ri[d];
        fu.v_out[3 * (size_t)i + d] = vi[d];
      }
      if (wrapped)
        for (int d = 0; d < 3; ++d) fu.mic_out[3 * (size_t)i + d] = mi[d];
      write_view(k, ss, fu.cell_of[i], ti, il, ri, vi, fu.pos4_out, fu.vel4_out);
      over = skin_one(k, ri, rr, mi);
    }
  }
  if (__builtin_amdgcn_ballot_w64(over) != 0ull && (threadIdx.x & 63) == 0)
    atomicOr(&fu.rebuild_flags[k.rep * F_NFLAGS + F_REBUILD], 1);
  zero_counts(fu.nc_tot, fu.cell_count);
}

template <int TPP, int XT, bool ENERGY, bool FUSE>
__global__ __launch_bounds__(MFA_BLOCK) void k_force64(const KParams* __restrict__ kb,
                                                       const int* __restrict__ rmap,
                                                       const double* __restrict__ r,
                                                       const double* __restrict__ v,
                                                       const int* __restrict__ type,
                                                       const int* __restrict__ perm,
                                                       const unsigned* __restrict__ nl,
                                                       const int* __restrict__ nn, int cap,
                                                       double* __restrict__ f,
                                                       const long long* __restrict__ d_step,
                                                       int mask, double* __restrict__ part,
                                                       int* __restrict__ flags_all, FuseArgs fu) {
  MFA_REPLICA(blockIdx.x * (MFA_BLOCK / TPP));
  int* flags = flags_all + k.rep * F_NFLAGS;
  __shared__ int err_s;
  // per type pair: range^2, epsilon, sigma^2, 1 / r_c of the DPD weight
  __shared__ double4 s_pair[MFA_NTYPE * MFA_NTYPE];
  if (threadIdx.x == 0) err_s = flags[F_ERROR];
  if (threadIdx.x >= 64 && threadIdx.x < 64 + MFA_NTYPE * MFA_NTYPE) {
    const int q = threadIdx.x - 64, a = q / MFA_NTYPE, b = q % MFA_NTYPE;
    s_pair[q] = make_double4(k.range_2[a][b], k.epsil[a][b], k.sigma_2[a][b],
                             k.dpd_cut_off_on ? k.inv_r_cut_dpd : k.inv_range[a][b]);
  }
  __syncthreads();
  if (err_s != 0) return;  // state already invalid: stop, the host reports it
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int sp = gtid / TPP;  // sorted position in the concatenated arrays
  const int sub = gtid % TPP;
  const bool active = sp - k.p0 < k.n;
  const int ss = active ? sp : k.p0;
  const int i = perm[ss];     // the particle (reference order)
  const int il = i - k.p0;    // within the replica (chains, heads, random keys)
  const int64_t step = *d_step;
  const double r2min = switch_on_r2min(k, step);
  const int ii = i;
  const double xi = r[3 * (size_t)ii + 0], yi = r[3 * (size_t)ii + 1], zi = r[3 * (size_t)ii + 2];
  const double vxi = v[3 * (size_t)ii + 0], vyi = v[3 * (size_t)ii + 1], vzi = v[3 * (size_t)ii + 2];
  const int ti = type[ii] - 1;
  double e_ff = 0.0, vir = 0.0, e_fw = 0.0, e_bond = 0.0, fzb = 0.0, fzt = 0.0, e_coul = 0.0;
  double e_bend = 0.0;
  double ax = 0.0, ay = 0.0, az = 0.0;

  if (active && (mask & TM_PAIR)) {
    const int wn = nn[ss], cnt = wn & 0xffff, nup = wn >> 16;
    const unsigned* row = nl + (size_t)ss * cap;
    // two candidates per lane and round: their list entries, then their
    // particles, then their positions and types are loaded together (the
    // order of the lane's sums is that of the entries)
    for (int q0 = sub; q0 < cnt; q0 += 2 * TPP) {
      int jv[2], tv[2];
      double xv[2], yv[2], zv[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int q = q0 + u * TPP;
        jv[u] = q < cnt ? perm[row[list_slot(q, nup, cnt, cap)] & MFA_JMASK] : ii;
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        xv[u] = r[3 * (size_t)jv[u] + 0];
        yv[u] = r[3 * (size_t)jv[u] + 1];
        zv[u] = r[3 * (size_t)jv[u] + 2];
        tv[u] = type[jv[u]];
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
      if (q0 + u * TPP >= cnt) continue;
      const int j = jv[u];
      double dx = xi - xv[u];
      double dy = yi - yv[u];
      double dz = zi - zv[u];
      min_image(k, dx, dy, dz);
      const double r2raw = (dx * dx + dy * dy) + dz * dz;
      const int tj = tv[u] - 1;
      const double4 tp = s_pair[ti * MFA_NTYPE + tj];
      const bool in_lj = r2raw < tp.x;
      // force switch-on clamp (fluid_fluid.f90:96); r2min = 0 leaves r2 as is.
      // The clamped value also feeds the DPD terms, as in the reference.
      const double r2 = in_lj ? fmax(r2raw, r2min) : r2raw;
      bool dpd = false;
      const double inv_rc = tp.w;  // 1 / r_c of the DPD weight
      if (k.dpd_on) {
        if (!k.dpd_cut_off_on) {
          dpd = in_lj;
        } else {
          const bool excl = k.dpd_head_excl && (ti == 0 || tj == 0);
          dpd = !excl && r2 < k.r_cut_dpd_2;
        }
      }
      // ewald_real (fluid_fluid.f90:166): inside the LJ range (compat) or rc_coul
      const bool coul = (XT & 1) && k.coulomb && (k.rc_coul_2 > 0.0 ? r2raw < k.rc_coul_2 : in_lj);
      if (!in_lj && !dpd && !coul) continue;
      double fx = 0.0, fy = 0.0, fz = 0.0;
      // 1/r from one reciprocal square root, r and 1/r^2 from it: each within
      // a few ulp of the reference's sqrt(r2), 1/r2 and sqrt(1/r2)
      const double inv_sqrt_r_2 = rsqrt(r2);
      const double inv_r_2 = inv_sqrt_r_2 * inv_sqrt_r_2;
      const double rr = r2 * inv_sqrt_r_2;
      if (in_lj) {
        const double eps = tp.y;
        const double r61 = tp.z * inv_r_2;
        const double r6 = (r61 * r61) * r61;
        const double r12 = r6 * r6;
        if (ENERGY) e_ff += 0.5 * (eps * ((r12 - r6) - k.e_shift[ti][tj]));
        const double rd = (eps * (-12.0 * r12 + 6.0 * r6)) * inv_r_2;
        fx = -(rd * dx);
        fy = -(rd * dy);
        fz = -(rd * dz);
        if (ENERGY) vir += 0.5 * (-(rd * r2raw));
      }
      if (dpd) {
        const double xi_n = pair_normal_f64(k.seed, step, il, j - k.p0);
        const double wr = 1.0 - rr * inv_rc;
        const double wd = wr * wr;
        const double ex = dx * inv_sqrt_r_2, ey = dy * inv_sqrt_r_2, ez = dz * inv_sqrt_r_2;
        const double sr = (k.sig * wr) * xi_n;
        fx = fx + sr * ex;
        fy = fy + sr * ey;
        fz = fz + sr * ez;
        const double dvx = vxi - v[3 * (size_t)j + 0];
        const double dvy = vyi - v[3 * (size_t)j + 1];
        const double dvz = vzi - v[3 * (size_t)j + 2];
        const double rv = (dvx * ex + dvy * ey) + dvz * ez;
        const double sd = ((-1.0 * k.friction1) * wd) * rv;
        fx = fx + sd * ex;
        fy = fy + sd * ey;
        fz = fz + sd * ez;
      }
      if (coul) {  // ewald_real.f90:15-32 with the clamped r_2
        const double qq = k.q_type[ti] * k.q_type[tj];
        const double inv_r = inv_sqrt_r_2;
        const double ec = erfc(k.sqrt_alpha * rr);
        if (ENERGY) e_coul += 0.5 * (qq * ec * inv_r);
        const double f_scal = qq * k.sqrt_alpha * k.inv_sqrt_pi * 2.0 * exp(-k.alpha_c * r2) * inv_r +
                              qq * ec * inv_r * inv_r;
        fx = fx + f_scal * dx * inv_r;
        fy = fy + f_scal * dy * inv_r;
        fz = fz + f_scal * dz * inv_r;
      }
      ax += fx;
      ay += fy;
      az += fz;
      }
    }
  }

  // Langevin, bonds and walls: lane 0 of the group (at most two bonds per particle).
  if ((XT & 2) && active && sub == 0 && k.lgv && (mask & TM_PAIR)) {
    double fl[3];
    langevin_force(k, step, il, ti, vxi, vyi, vzi, fl);
    ax += fl[0];
    ay += fl[1];
    az += fl[2];
  }
  if (active && sub == 0) {
    if (mask & TM_BOND) {
      int base, len;
      chain_of(k, il, base, len);
      for (int side = 0; side < 2; ++side) {
        // bond (lo, lo+1): side 0 -> (i-1, i), side 1 -> (i, i+1)
        const int lol = side == 0 ? il - 1 : il;
        if (len < 2 || lol < base || lol + 1 >= base + len) continue;
        const int lo = k.p0 + lol, hi = lo + 1;
        double dx = r[3 * (size_t)lo + 0] - r[3 * (size_t)hi + 0];
        double dy = r[3 * (size_t)lo + 1] - r[3 * (size_t)hi + 1];
        double dz = r[3 * (size_t)lo + 2] - r[3 * (size_t)hi + 2];
        min_image(k, dx, dy, dz);
        const double r2 = (dx * dx + dy * dy) + dz * dz;
        double x = r2 * k.inv_r_chain_2;
        const double clamp = (double)0.98f;
        x = x < clamp ? x : clamp;
        const double rd = (k.k_chain / (1.0 - x)) * k.inv_r_chain_2;
        const double sgn = side == 0 ? 1.0 : -1.0;  // +force_loc on hi, -force_loc on lo
        ax += sgn * (rd * dx);
        ay += sgn * (rd * dy);
        az += sgn * (rd * dz);
        if (ENERGY && side == 1) {
          e_bond += -0.5 * log(1.0 - x);
          vir += -(rd * r2);
        }
      }
    }
    if ((XT & 1) && (mask & TM_BEND) && k.bending && il < k.n_brush) {
      const int base = (il / k.n_mon) * k.n_mon, m = il - base;
      for (int cen = m - 1; cen <= m + 1; ++cen) {
        if (cen < 1 || cen > k.n_mon - 2) continue;
        double fp[3], fn[3];
        const double eb = bend_angle(k, r, k.p0 + base + cen, fp, fn);
        if (cen == m) {
          e_bend += eb;
          ax += -fp[0] - fn[0];
          ay += -fp[1] - fn[1];
          az += -fp[2] - fn[2];
        } else if (cen == m + 1) {
          ax += fp[0];
          ay += fp[1];
          az += fp[2];
        } else {
          ax += fn[0];
          ay += fn[1];
          az += fn[2];
        }
      }
    }
    if ((mask & TM_WALL) && k.symmetry == 0) {
      wall_93(k, ti, zi, fzb, fzt, e_fw);
      az += fzb;
      az += fzt;
    }
  }

  // Combine the lane group and write F_i.
  double fx = group_sum<TPP>(ax), fy = group_sum<TPP>(ay), fz = group_sum<TPP>(az);
  if (FUSE) {
    if (active && sub == 0 && !(isfinite(fx) && isfinite(fy) && isfinite(fz)))
      atomicCAS(&flags[F_ERROR], 0, MFA_ERR_NAN);
  } else if (active && sub == 0) {
    if (mask & TM_ACCUM) {
      fx = f[3 * (size_t)i + 0] + fx;
      fy = f[3 * (size_t)i + 1] + fy;
      fz = f[3 * (size_t)i + 2] + fz;
    }
    f[3 * (size_t)i + 0] = fx;
    f[3 * (size_t)i + 1] = fy;
    f[3 * (size_t)i + 2] = fz;
    if (!(isfinite(fx) && isfinite(fy) && isfinite(fz))) atomicCAS(&flags[F_ERROR], 0, MFA_ERR_NAN);
  }
  if (ENERGY) {
    double vals[MFA_NTERMS];
    for (int t = 0; t < MFA_NTERMS; ++t) vals[t] = 0.0;
    if (active) {
      vals[T_VFF] = e_ff;
      vals[T_VIR] = vir;
      vals[T_VFW] = e_fw;
      vals[T_VINTRA] = e_bond * k.k_chain;
      vals[T_VCOUL] = e_coul;
      vals[T_VBEND] = e_bend;
      vals[T_FZB] = fzb;
      vals[T_FZT] = fzt;
      if (sub == 0 && ti == 1) vals[T_ZMOM] = zi;
    }
    block_partials(vals, part);
  }
  // a step's pass: the k-space force (charged systems), verlet_velocities of
  // this step, verlet_positions and check_skin of the next
  if (FUSE) fused_epilogue<TPP, (XT & 1) != 0, false>(k, active, sub, i, il, ti, ss, fx, fy, fz, r, fu);
}

// The scalar s of one interacting pair of the mixed form, F_ij = s (dx, dy,
// dz): LJ with the switch-on clamp (skipped for epsilon 0: there it is
// exactly 0 in FP64, while in FP32 r^-12 r^-2 overflows below r ~ 0.002,
// where such beads may sit, and 0 * inf would be NaN), DPD and the
// real-space Coulomb term as the flags say (fluid_fluid.f90:94-126,
// dpd_forces.f90:19-94, ewald_real.f90:15-32); with ENERGY the pair's half
// of the energies and virial.
template <int XT, bool ENERGY>
__device__ __forceinline__ float pair_scalar(const KParams& k, const float4 tb, float inv_rc,
                                             bool in_lj, bool dpd, bool coul, float r2raw,
                                             float r2min, float dx, float dy, float dz,
                                             const float4 vi4, const float4 vj,
                                             uint32_t seed, int64_t step, int il, int jl,
                                             int ti, int tj, float sigf, float gam, float sa,
                                             float alf, float pref, double& e_ff, double& vir,
                                             double& e_coul) {
#pragma clang fp contract(fast)
  const float r2 = in_lj ? fmaxf(r2raw, r2min) : r2raw;
  const float inv_r = __builtin_amdgcn_rsqf(r2);  // v_rsq_f32, ~1 ulp
  const float inv_r_2 = inv_r * inv_r;
  float s = 0.0f;
  if (in_lj && tb.y != 0.0f) {
    const float r61 = tb.z * inv_r_2;
    const float r6 = r61 * r61 * r61;
    const float r12 = r6 * r6;
    const float rd = tb.y * (6.0f * r6 - 12.0f * r12) * inv_r_2;
    s = -rd;
    if (ENERGY) {
      e_ff += 0.5 * (double)(tb.y * ((r12 - r6) - tb.w));
      vir += 0.5 * (double)(-rd * r2raw);
    }
  }
  if (dpd) {
    const float xi_n = pair_normal_f32(seed, step, il, jl);
    // w_r = 1 - r / rc with r = r^2 / r (the rsq above)
    const float wr = 1.0f - (r2 * inv_r) * inv_rc;
    const float rv = ((vi4.x - vj.x) * dx + (vi4.y - vj.y) * dy + (vi4.z - vj.z) * dz) * inv_r;
    s += (sigf * wr * xi_n - gam * (wr * wr) * rv) * inv_r;
  }
  if ((XT & 1) && coul) {  // ewald_real.f90:15-32 with the clamped r_2
    const float qq = (float)k.q_type[ti] * (float)k.q_type[tj];
    const float rr = r2 * inv_r;
    const float ec = erfcf(sa * rr);
    s += (qq * pref * __expf(-alf * r2) * inv_r + qq * ec * inv_r_2) * inv_r;
    if (ENERGY) e_coul += 0.5 * (double)(qq * ec * inv_r);
  }
  return s;
}

// Candidates per lane whose loads the mixed force pass issues together: 2
// below 16 lanes per particle (72 VGPRs, 7 waves per SIMD: the melt's
// 6,250 waves at 4 lanes fit the chip in one round), 4 at 16 lanes (waves
// enough at 5 per SIMD; more loads in flight per wave).
template <int TPP>
struct FmBatch {
  static constexpr int value = TPP >= 16 ? 4 : 2;
};
// Pairs a wave's queue holds (a round adds at most 64 to fewer than 64).
#define FMQ_CAP 128

// XT bits: 1 Coulomb real space and bending, 2 the Langevin force, 4 DPD_CUT_OFF
// (a DPD range of its own, heads excluded from DPD).  MODE: 0 forces, 1 forces
// and energy partials, 2 forces, then verlet_velocities of this step and
// verlet_positions and check_skin of the next (k_fused_integrate's work) in
// the same pass, which leaves F unwritten; 3 the first pass of the half-list
// form: the upper entries of each row only, each interacting pair's scaled
// scalar kept in hf.pair_s (0 for the others), the particle's terms
// (pairs, bonds, walls, Langevin, bending) left in hf.fpart as fixed-point
// sums for k_force_lower; after a rebuild it also finds, for each lower
// entry, the slot of its pair in the neighbour's row (hf.mir).
template <int TPP, int MODE, int XT, int QM>
__global__ __launch_bounds__(MFA_BLOCK) void k_force_mixed(
    const KParams* __restrict__ kb, const int* __restrict__ rmap, const double* __restrict__ r,
    const float4* __restrict__ pos4, const float4* __restrict__ vel4, const int* __restrict__ perm,
    const unsigned* __restrict__ nl, const int* __restrict__ nn, int cap, double* __restrict__ f,
    const long long* __restrict__ d_step, int mask, double* __restrict__ part,
    int* __restrict__ flags_all, FuseArgs fu, HalfArgs hf) {
#pragma clang fp contract(fast)
  constexpr bool ENERGY = MODE == 1;
  constexpr bool FUSE = MODE == 2;
  constexpr bool HALF = MODE == 3;
  constexpr bool DC = (XT & 4) != 0;
  constexpr int FM_BATCH = FmBatch<TPP>::value;
  MFA_REPLICA(blockIdx.x * (MFA_BLOCK / TPP));
  int* flags = flags_all + k.rep * F_NFLAGS;
  // cell offsets -(d)*r_bin of the 27 codes, and the FP32 pair tables
  __shared__ float4 s_off[27];
  __shared__ float4 s_tab[MFA_NTYPE * MFA_NTYPE];  // range_2, 4 eps, sigma_2, e_shift
  __shared__ float s_inv_rc[MFA_NTYPE * MFA_NTYPE];  // 1 / sqrt(range_2)
  __shared__ int err_s, mirror_s;
  if (threadIdx.x < 27) {
    const int c = threadIdx.x;
    const int dx = c / 9 - 1, dy = (c / 3) % 3 - 1, dz = c % 3 - 1;
    s_off[c] = make_float4(-(float)dx * (float)k.r_bin[0], -(float)dy * (float)k.r_bin[1],
                           -(float)dz * (float)k.r_bin[2], 0.0f);
  }
  if (threadIdx.x >= 32 && threadIdx.x < 32 + MFA_NTYPE * MFA_NTYPE) {
    const int q = threadIdx.x - 32, a = q / MFA_NTYPE, b = q % MFA_NTYPE;
    s_tab[q] = make_float4((float)k.range_2[a][b], (float)k.epsil[a][b], (float)k.sigma_2[a][b],
                           (float)k.e_shift[a][b]);
    s_inv_rc[q] = (float)sqrt(k.inv_range_2[a][b]);
  }
  if (threadIdx.x == 64) err_s = flags[F_ERROR];
  if (threadIdx.x == 65) mirror_s = HALF ? flags[F_MIRROR] : 0;
  __syncthreads();
  if (err_s != 0) return;  // state already invalid: stop, the host reports it
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int sp = gtid / TPP;  // sorted position in the concatenated arrays
  const int sub = gtid % TPP;
  const bool active = sp - k.p0 < k.n;
  const int ss = active ? sp : k.p0;
  const int i = perm[ss];     // the particle (reference order)
  const int il = i - k.p0;    // within the replica (chains, heads, random keys)
  const int64_t step = *d_step;
  const double r2min_d = switch_on_r2min(k, step);
  const float r2min = (float)r2min_d;
  const float rdpd2 = (float)k.r_cut_dpd_2, inv_rdpd = (float)(1.0 / sqrt(k.r_cut_dpd_2));
  // DPD_CUT_OFF under the switch-on clamp: the reference tests the clamped
  // r^2 of a pair inside the LJ range against the DPD range
  // (fluid_fluid.f90:94-96,201), so no such pair has a DPD term while
  // r2min >= r_cut_dpd^2 (decided in FP64, as the clamp is)
  const bool lj_no_dpd = DC && r2min_d >= k.r_cut_dpd_2;
  const float sigf = (float)k.sig, gam = (float)k.friction1, margin = k.ff_margin;
  const float4 pi4 = pos4[ss];
  const float4 vi4 = vel4[ss];
  const int ti = __float_as_int(pi4.w) & 3;
  double e_ff = 0.0, vir = 0.0, e_fw = 0.0, e_bond = 0.0, fzb = 0.0, fzt = 0.0, zi = 0.0;
  double e_coul = 0.0, e_bend = 0.0;
  const float rc2f = (float)k.rc_coul_2, sa = (float)k.sqrt_alpha, alf = (float)k.alpha_c;
  const float pref = (float)(k.sqrt_alpha * k.inv_sqrt_pi * 2.0);
  double ax = 0.0, ay = 0.0, az = 0.0;  // fixed-point sums (integer-valued)

  // run-time switches uniform over the launch, read once
  const uint32_t seed = k.seed;
  const bool dpd_on = k.dpd_on != 0;
  const bool head_excl = DC && k.dpd_head_excl != 0;
  const bool coul_on = (XT & 1) && k.coulomb != 0;
  const bool coul_rc = (XT & 1) && k.rc_coul_2 > 0.0;  // own real-space cut-off
  bool qbad = false;  // a queued pair's fixed-point terms out of range
  // A row's entries lie in two parts (list_slot): the front [0, n_front)
  // and the back [cap - n_back, cap).  Near/shell rows hold in front the
  // entries inside the interaction range at the build, in the back those in
  // the skin; half-list rows hold in front the upper entries (the half-list
  // first pass walks only those); plain rows have everything in front.  QM
  // says how each part is walked: 0 both directly, 1 both through the pair
  // queue, 2 the front directly and the back through the queue.
  const int wn = active && (mask & TM_PAIR) ? nn[ss] : 0;
  const int n_front = wn >> 16, n_back = HALF ? 0 : (wn & 0xffff) - (wn >> 16);
  const unsigned* row = nl + (size_t)ss * cap;

  // Walks entries rw[0, cnt) directly (their row slots slot0 + q).  FM_BATCH
  // candidates per lane and round: their list entries are loaded together,
  // then their views together, so each round waits for two memory round
  // trips instead of two per candidate.  Past the end a slot reads the
  // particle's own view and is skipped.
  auto direct_range = [&](const unsigned* rw, int cnt, int slot0) {
    float* srow = HALF ? hf.pair_s + (size_t)ss * cap + slot0 : nullptr;
    for (int q0 = sub; q0 < cnt; q0 += TPP * FM_BATCH) {
      unsigned ev[FM_BATCH];
      float4 pv[FM_BATCH];
#pragma unroll
      for (int u = 0; u < FM_BATCH; ++u) {
        const int q = q0 + u * TPP;
        ev[u] = q < cnt ? rw[q] : (unsigned)ss;
      }
#pragma unroll
      for (int u = 0; u < FM_BATCH; ++u) pv[u] = pos4[(int)(ev[u] & MFA_JMASK)];
      // the candidates' velocities with their positions (the DPD term needs
      // them for each interacting pair), so their latency is not exposed
      // inside the pair term
      float4 vv[FM_BATCH];
#pragma unroll
      for (int u = 0; u < FM_BATCH; ++u)
        vv[u] = dpd_on ? vel4[(int)(ev[u] & MFA_JMASK)] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
      for (int u = 0; u < FM_BATCH; ++u) {
        if (q0 + u * TPP >= cnt) continue;
        const unsigned e = ev[u];
        const float4 o = s_off[e >> MFA_JBITS];
        const float4 pj = pv[u];
        const float dx = (pi4.x - pj.x) + o.x;
        const float dy = (pi4.y - pj.y) + o.y;
        const float dz = (pi4.z - pj.z) + o.z;
        const float r2raw = dx * dx + dy * dy + dz * dz;
        const int wj = __float_as_int(pj.w);
        const int tj = wj & 3, jl = wj >> 2;  // type and id within the replica
        const float4 tb = s_tab[ti * MFA_NTYPE + tj];
        bool in_lj = r2raw < tb.x;
        bool in_dpd = DC ? r2raw < rdpd2 && !(in_lj && lj_no_dpd) : in_lj;
        bool in_coul = coul_rc && r2raw < rc2f;
        // distance of r^2 to the nearest cut-off that decides something
        float dm = fabsf(r2raw - tb.x);
        if (DC) dm = fminf(dm, fabsf(r2raw - rdpd2));
        if (coul_rc) dm = fminf(dm, fabsf(r2raw - rc2f));
        if (__builtin_expect(dm < margin, 0)) {
          // within the FP32 error of a cut-off: decide in FP64 (fluid_fluid.f90:84-94)
          const int j = k.p0 + jl;
          double ex = r[3 * (size_t)i + 0] - r[3 * (size_t)j + 0];
          double ey = r[3 * (size_t)i + 1] - r[3 * (size_t)j + 1];
          double ez = r[3 * (size_t)i + 2] - r[3 * (size_t)j + 2];
          min_image(k, ex, ey, ez);
          const double r2d = (ex * ex + ey * ey) + ez * ez;
          in_lj = r2d < k.range_2[ti][tj];
          in_dpd = DC ? r2d < k.r_cut_dpd_2 && !(in_lj && lj_no_dpd) : in_lj;
          in_coul = coul_rc && r2d < k.rc_coul_2;
        }
        // DPD_CUT_OFF excludes the heads (fluid_fluid.f90:179-185)
        const bool dpd = dpd_on && in_dpd && !(head_excl && (ti == 0 || tj == 0));
        // ewald_real (fluid_fluid.f90:166): inside the LJ range (compat) or rc_coul
        const bool coul = coul_on && (coul_rc ? in_coul : in_lj);
        if (!in_lj && !dpd && !coul) {
          if (HALF) srow[q0 + u * TPP] = 0.0f;
          continue;
        }
        const float s = pair_scalar<XT, ENERGY>(
            k, tb, DC ? inv_rdpd : s_inv_rc[ti * MFA_NTYPE + tj], in_lj, dpd, coul, r2raw, r2min, dx,
            dy, dz, vi4, vv[u], seed, step, il, jl, ti, tj, sigf, gam, sa, alf, pref, e_ff, vir,
            e_coul);
        // the pair's three components in fixed point: (s 2^22) d equals
        // (s d) 2^22 exactly (a power-of-two scale), truncated toward zero
        const float sf = s * MFA_FIX_SCALE;
        if (HALF) srow[q0 + u * TPP] = sf;
        ax += (double)truncf(sf * dx);
        ay += (double)truncf(sf * dy);
        az += (double)truncf(sf * dz);
      }
    }
  };

  if constexpr (QM != 0) {
    // The pair queue: the same pairs, with the wave's work compacted.  Each
    // lane tests its candidates (as the direct walk does) and pushes those
    // with a term to evaluate into the wave's queue in LDS; whenever 64 are
    // queued (and at the end of a part) the whole wave evaluates them, one
    // pair per lane, and adds each pair's three fixed-point terms to its
    // particle's sums in LDS with 64-bit integer atomics.  The terms are the
    // direct walk's, and integer sums do not depend on order, so the forces
    // equal the direct walk's bit for bit; only the lanes are kept busy where
    // few candidates interact (a DPD cut-off below the list radius, the skin
    // part of near/shell rows).  Everything here is per wave, in
    // wave-uniform control flow.
    constexpr int NS = 64 / TPP;  // particles (lane groups) per wave
    __shared__ int4 q_rec[MFA_BLOCK / 64][FMQ_CAP][2];
    __shared__ float4 q_vi[MFA_BLOCK / 64][NS];
    __shared__ int q_id[MFA_BLOCK / 64][NS];  // (il << 2) | ti
    __shared__ unsigned long long q_acc[MFA_BLOCK / 64][NS][3];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, slot = lane / TPP;
    if (sub == 0) {
      q_vi[wv][slot] = vi4;
      q_id[wv][slot] = (il << 2) | ti;
      q_acc[wv][slot][0] = 0ull;
      q_acc[wv][slot][1] = 0ull;
      q_acc[wv][slot][2] = 0ull;
    }
    int head = 0, tail = 0;  // queue positions, the same in every lane
    // evaluates queued pairs head .. head + n - 1, one per lane
    auto drain = [&](int n) {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (lane < n) {
        const int at = (head + lane) & (FMQ_CAP - 1);
        const int4 a = q_rec[wv][at][0], b = q_rec[wv][at][1];
        const float qdx = __int_as_float(a.x), qdy = __int_as_float(a.y), qdz = __int_as_float(a.z);
        const int sl = b.z & 0xff, qtj = (b.z >> 8) & 3;
        const int idq = q_id[wv][sl];
        const int qti = idq & 3;
        const float4 qvj = (b.z >> 11) & 1 ? vel4[b.x] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float qs = pair_scalar<XT, ENERGY>(
            k, s_tab[qti * MFA_NTYPE + qtj], DC ? inv_rdpd : s_inv_rc[qti * MFA_NTYPE + qtj],
            (b.z >> 10) & 1, (b.z >> 11) & 1, (b.z >> 12) & 1, __int_as_float(a.w), r2min, qdx, qdy,
            qdz, q_vi[wv][sl], qvj, seed, step, idq >> 2, b.y, qti, qtj, sigf, gam, sa, alf,
            pref, e_ff, vir, e_coul);
        const float sf = qs * MFA_FIX_SCALE;
        if (HALF) hf.pair_s[(unsigned)b.w] = sf;  // the pair's slot (row * cap + entry)
        const float cx = truncf(sf * qdx), cy = truncf(sf * qdy), cz = truncf(sf * qdz);
        // |term| < 2^52, NaN excluded: then the int64 sums are exact
        if (fabsf(cx) < 4.5e15f && fabsf(cy) < 4.5e15f && fabsf(cz) < 4.5e15f) {
          atomicAdd(&q_acc[wv][sl][0], (unsigned long long)(long long)cx);
          atomicAdd(&q_acc[wv][sl][1], (unsigned long long)(long long)cy);
          atomicAdd(&q_acc[wv][sl][2], (unsigned long long)(long long)cz);
        } else {
          qbad = true;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    };
    // Tests entries rw[0, cnt) (row slots slot0 + q) and queues the pairs
    // with a term; every lane of the wave calls it.
    auto queue_range = [&](const unsigned* rw, int cnt, int slot0) {
      int mcnt = cnt;
      for (int o = 32; o > 0; o >>= 1) mcnt = max(mcnt, __shfl_xor(mcnt, o, 64));
      mcnt = __builtin_amdgcn_readfirstlane(mcnt);
      for (int q0 = 0; q0 < mcnt; q0 += TPP * FM_BATCH) {
        unsigned ev[FM_BATCH];
        float4 pv[FM_BATCH];
#pragma unroll
        for (int u = 0; u < FM_BATCH; ++u) {
          const int q = q0 + sub + u * TPP;
          ev[u] = q < cnt ? rw[q] : (unsigned)ss;
        }
#pragma unroll
        for (int u = 0; u < FM_BATCH; ++u) pv[u] = pos4[(int)(ev[u] & MFA_JMASK)];
#pragma unroll
        for (int u = 0; u < FM_BATCH; ++u) {
          const bool valid = q0 + sub + u * TPP < cnt;
          const unsigned e = ev[u];
          const int js = (int)(e & MFA_JMASK);
          const float4 o = s_off[e >> MFA_JBITS];
          const float4 pj = pv[u];
          const float dx = (pi4.x - pj.x) + o.x;
          const float dy = (pi4.y - pj.y) + o.y;
          const float dz = (pi4.z - pj.z) + o.z;
          const float r2raw = dx * dx + dy * dy + dz * dz;
          const int wj = __float_as_int(pj.w);
          const int tj = wj & 3, jl = wj >> 2;
          const float4 tb = s_tab[ti * MFA_NTYPE + tj];
          bool in_lj = r2raw < tb.x;
          bool in_dpd = DC ? r2raw < rdpd2 && !(in_lj && lj_no_dpd) : in_lj;
          bool in_coul = coul_rc && r2raw < rc2f;
          float dm = fabsf(r2raw - tb.x);
          if (DC) dm = fminf(dm, fabsf(r2raw - rdpd2));
          if (coul_rc) dm = fminf(dm, fabsf(r2raw - rc2f));
          if (__builtin_expect(valid && dm < margin, 0)) {
            // within the FP32 error of a cut-off: decide in FP64 (fluid_fluid.f90:84-94)
            const int j = k.p0 + jl;
            double ex = r[3 * (size_t)i + 0] - r[3 * (size_t)j + 0];
            double ey = r[3 * (size_t)i + 1] - r[3 * (size_t)j + 1];
            double ez = r[3 * (size_t)i + 2] - r[3 * (size_t)j + 2];
            min_image(k, ex, ey, ez);
            const double r2d = (ex * ex + ey * ey) + ez * ez;
            in_lj = r2d < k.range_2[ti][tj];
            in_dpd = DC ? r2d < k.r_cut_dpd_2 && !(in_lj && lj_no_dpd) : in_lj;
            in_coul = coul_rc && r2d < k.rc_coul_2;
          }
          const bool dpd = dpd_on && in_dpd && !(head_excl && (ti == 0 || tj == 0));
          const bool coul = coul_on && (coul_rc ? in_coul : in_lj);
          const bool push = valid && ((in_lj && tb.y != 0.0f) || dpd || coul);
          if (HALF && valid && !push)
            hf.pair_s[(size_t)ss * cap + (slot0 + q0 + sub + u * TPP)] = 0.0f;
          const unsigned long long pm = __builtin_amdgcn_ballot_w64(push);
          if (push) {
            const int at = (tail + (int)__builtin_amdgcn_mbcnt_hi(
                                       (unsigned)(pm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)pm, 0u))) &
                           (FMQ_CAP - 1);
            q_rec[wv][at][0] = make_int4(__float_as_int(dx), __float_as_int(dy), __float_as_int(dz),
                                         __float_as_int(r2raw));
            q_rec[wv][at][1] = make_int4(js, jl,
                                         slot | (tj << 8) | ((in_lj ? 1 : 0) << 10) |
                                             ((dpd ? 1 : 0) << 11) | ((coul ? 1 : 0) << 12),
                                         HALF ? ss * cap + (slot0 + q0 + sub + u * TPP) : 0);
          }
          tail += (int)__popcll(pm);
          if (tail - head >= 64) {
            drain(64);
            head += 64;
          }
        }
      }
      while (tail > head) {
        const int n = tail - head < 64 ? tail - head : 64;
        drain(n);
        head += n;
      }
    };
    if (QM == 1) queue_range(row, n_front, 0);
    else direct_range(row, n_front, 0);
    queue_range(row + (cap - n_back), n_back, cap - n_back);
    if (sub == 0) {
      ax += (double)(long long)q_acc[wv][slot][0];
      ay += (double)(long long)q_acc[wv][slot][1];
      az += (double)(long long)q_acc[wv][slot][2];
    }
  } else {
    direct_range(row, n_front, 0);
    direct_range(row + (cap - n_back), n_back, cap - n_back);
  }

  // Langevin, bonds and walls: lane 0 of the group, from the FP64 positions.
  if ((XT & 2) && active && sub == 0 && k.lgv && (mask & TM_PAIR)) {
    double fl[3];
    langevin_force(k, step, il, ti, (double)vi4.x, (double)vi4.y, (double)vi4.z, fl);
    fix_add3(ax, ay, az, (float)fl[0], (float)fl[1], (float)fl[2]);
  }
  if (active && sub == 0 && (ENERGY || (mask & (TM_BOND | TM_WALL)))) zi = r[3 * (size_t)i + 2];
  if ((XT & 1) && active && sub == 0 && (mask & TM_BEND) && k.bending && il < k.n_brush) {
    // each term through float and fixed point; the centre takes minus the
    // fixed-point sum of its neighbours' terms, so every angle sums to zero
    const int base = (il / k.n_mon) * k.n_mon, m = il - base;
    for (int cen = m - 1; cen <= m + 1; ++cen) {
      if (cen < 1 || cen > k.n_mon - 2) continue;
      double fp[3], fn[3];
      const double eb = bend_angle(k, r, k.p0 + base + cen, fp, fn);
      double px = 0.0, py = 0.0, pz = 0.0, nx = 0.0, ny = 0.0, nz = 0.0;
      fix_add3(px, py, pz, (float)fp[0], (float)fp[1], (float)fp[2]);
      fix_add3(nx, ny, nz, (float)fn[0], (float)fn[1], (float)fn[2]);
      if (cen == m) {
        if (ENERGY) e_bend += eb;
        ax -= px + nx;
        ay -= py + ny;
        az -= pz + nz;
      } else if (cen == m + 1) {
        ax += px;
        ay += py;
        az += pz;
      } else {
        ax += nx;
        ay += ny;
        az += nz;
      }
    }
  }
  if (active && sub == 0 && (mask & (TM_BOND | TM_WALL))) {
    if (mask & TM_BOND) {
      int base, len;
      chain_of(k, il, base, len);
      for (int side = 0; side < 2; ++side) {
        const int lol = side == 0 ? il - 1 : il;
        if (len < 2 || lol < base || lol + 1 >= base + len) continue;
        const int lo = k.p0 + lol, hi = lo + 1;
        double dx = r[3 * (size_t)lo + 0] - r[3 * (size_t)hi + 0];
        double dy = r[3 * (size_t)lo + 1] - r[3 * (size_t)hi + 1];
        double dz = r[3 * (size_t)lo + 2] - r[3 * (size_t)hi + 2];
        min_image(k, dx, dy, dz);
        const double r2 = (dx * dx + dy * dy) + dz * dz;
        float x = (float)(r2 * k.inv_r_chain_2);
        x = x < 0.98f ? x : 0.98f;
        const float rd = (float)k.k_chain / (1.0f - x) * (float)k.inv_r_chain_2;
        const float sg = side == 0 ? rd : -rd;
        fix_add3(ax, ay, az, sg * (float)dx, sg * (float)dy, sg * (float)dz);
        if (ENERGY && side == 1) {
          e_bond += -0.5 * log(1.0 - (double)x);
          vir += (double)(-rd * (float)r2);
        }
      }
    }
    if ((mask & TM_WALL) && k.symmetry == 0) {
      wall_93(k, ti, zi, fzb, fzt, e_fw);
      fix_add3(ax, ay, az, 0.0f, 0.0f, (float)fzb);
      fix_add3(ax, ay, az, 0.0f, 0.0f, (float)fzt);
    }
  }

  if (HALF) {
    // After a rebuild: each lower entry (i -> j, j below i in sorted order)
    // finds its pair's slot in j's row, where i is an upper entry with the
    // opposite cell offset code; the upper part of a row ascends in entry
    // value, so a bisection over it finds the slot.
    if (mirror_s != 0 && active) {
      const int wn = nn[ss], tot = wn & 0xffff, nlo = tot - (wn >> 16);
      const unsigned* row = nl + (size_t)ss * cap;
      for (int q = sub; q < nlo; q += TPP) {
        const int slot = cap - nlo + q;
        const unsigned e = row[slot];
        const int a = (int)(e & MFA_JMASK);
        const unsigned key = (unsigned)ss | ((26u - (e >> MFA_JBITS)) << MFA_JBITS);
        const unsigned* ra = nl + (size_t)a * cap;
        const int na = nn[a] >> 16;
        int lo = 0, hi = na;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (ra[mid] < key) lo = mid + 1;
          else hi = mid;
        }
        if (lo >= na || ra[lo] != key) atomicCAS(&flags[F_ERROR], 0, MFA_ERR_INTERNAL);
        hf.mir[(size_t)ss * cap + slot] = (uint16_t)(lo < na ? lo : 0);
      }
    }
    // the group's fixed-point sums, for the second pass
    const bool bad = fix_bad(ax, ay, az) || qbad;
    const double sx = group_sum<TPP>(ax), sy = group_sum<TPP>(ay), sz = group_sum<TPP>(az);
    if (active && sub == 0) {
      hf.fpart[3 * (size_t)ss + 0] = sx;
      hf.fpart[3 * (size_t)ss + 1] = sy;
      hf.fpart[3 * (size_t)ss + 2] = sz;
    }
    if (bad) atomicCAS(&flags[F_ERROR], 0, MFA_ERR_NAN);
    return;
  }

  // Combine the lane group and write F_i.
  const bool bad = fix_bad(ax, ay, az) || qbad;
  double fx = group_sum<TPP>(ax) * MFA_FIX_INV;
  double fy = group_sum<TPP>(ay) * MFA_FIX_INV;
  double fz = group_sum<TPP>(az) * MFA_FIX_INV;
  if (active && sub == 0 && !FUSE) {
    if (mask & TM_ACCUM) {
      fx = f[3 * (size_t)i + 0] + fx;
      fy = f[3 * (size_t)i + 1] + fy;
      fz = f[3 * (size_t)i + 2] + fz;
    }
    f[3 * (size_t)i + 0] = fx;
    f[3 * (size_t)i + 1] = fy;
    f[3 * (size_t)i + 2] = fz;
  }
  if (bad) atomicCAS(&flags[F_ERROR], 0, MFA_ERR_NAN);
  if (ENERGY) {
    double vals[MFA_NTERMS];
    for (int t = 0; t < MFA_NTERMS; ++t) vals[t] = 0.0;
    // pair terms from every lane: with the pair queue a lane beyond the
    // replica's particles may evaluate another lane's pairs (zero otherwise)
    vals[T_VFF] = e_ff;
    vals[T_VIR] = vir;
    vals[T_VCOUL] = e_coul;
    if (active) {
      vals[T_VFW] = e_fw;
      vals[T_VINTRA] = e_bond * k.k_chain;
      vals[T_VBEND] = e_bend;
      vals[T_FZB] = fzb;
      vals[T_FZT] = fzt;
      if (sub == 0 && ti == 1) vals[T_ZMOM] = zi;
    }
    block_partials(vals, part);
  }
  // (charged systems carry XT bit 1; only they add a k-space force)
  if (FUSE) fused_epilogue<TPP, (XT & 1) != 0, true>(k, active, sub, i, il, ti, ss, fx, fy, fz, r, fu);
}

// Second pass of the half-list form.  Each lower entry of a row (i -> j, j
// below i in sorted order) takes its pair's scaled scalar from the slot of
// j's row that the first pass wrote for the same pair (found after each
// rebuild, hf.mir) and forms its fixed-point terms from its own separation,
// which is the negated separation j's row used (the FP32 view differences
// and cell offsets negate exactly); truncation toward zero is odd, so each
// term is exactly minus the other row's and F_i equals the full-list pass's
// bit for bit.  Then the first pass's sums of i are added and F_i written
// (MODE 0, as k_force_mixed MODE 0 writes it) or the step integrated (MODE
// 2, as k_force_mixed MODE 2).  The pass also clears the replica's F_MIRROR:
// the first pass, which reads it, has finished.
template <int TPP, int MODE>
__global__ __launch_bounds__(MFA_BLOCK) void k_force_lower(
    const KParams* __restrict__ kb, const int* __restrict__ rmap, const double* __restrict__ r,
    const float4* __restrict__ pos4, const int* __restrict__ perm, const unsigned* __restrict__ nl,
    const int* __restrict__ nn, int cap, double* __restrict__ f, int mask,
    int* __restrict__ flags_all, FuseArgs fu, HalfArgs hf) {
#pragma clang fp contract(fast)
  constexpr bool FUSE = MODE == 2;
  constexpr int FM_BATCH = 4;  // slots per lane whose loads are issued together
  const int sp0 = blockIdx.x * (MFA_BLOCK / TPP);
  MFA_REPLICA(sp0);
  int* flags = flags_all + k.rep * F_NFLAGS;
  __shared__ float4 s_off[27];
  __shared__ int err_s;
  if (threadIdx.x < 27) {
    const int c = threadIdx.x;
    const int dx = c / 9 - 1, dy = (c / 3) % 3 - 1, dz = c % 3 - 1;
    s_off[c] = make_float4(-(float)dx * (float)k.r_bin[0], -(float)dy * (float)k.r_bin[1],
                           -(float)dz * (float)k.r_bin[2], 0.0f);
  }
  if (threadIdx.x == 64) err_s = flags[F_ERROR];
  __syncthreads();
  if (sp0 == k.p0 && threadIdx.x == 0) flags[F_MIRROR] = 0;
  if (err_s != 0) return;
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int sp = gtid / TPP, sub = gtid % TPP;
  const bool active = sp - k.p0 < k.n;
  const int ss = active ? sp : k.p0;
  const int i = perm[ss];
  const int il = i - k.p0;
  const float4 pi4 = pos4[ss];
  const int ti = __float_as_int(pi4.w) & 3;
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (active && (mask & TM_PAIR)) {
    const int wn = nn[ss], nlo = (wn & 0xffff) - (wn >> 16);
    const size_t base = (size_t)ss * cap + (size_t)(cap - nlo);
    for (int q0 = sub; q0 < nlo; q0 += TPP * FM_BATCH) {
      unsigned ev[FM_BATCH];
      int mv[FM_BATCH];
      float sv[FM_BATCH];
#pragma unroll
      for (int u = 0; u < FM_BATCH; ++u) {
        const int q = q0 + u * TPP;
        ev[u] = q < nlo ? nl[base + q] : (unsigned)ss;
        mv[u] = q < nlo ? (int)hf.mir[base + q] : -1;
      }
#pragma unroll
      for (int u = 0; u < FM_BATCH; ++u)
        sv[u] = mv[u] >= 0 ? hf.pair_s[(size_t)(ev[u] & MFA_JMASK) * cap + mv[u]] : 0.0f;
#pragma unroll
      for (int u = 0; u < FM_BATCH; ++u) {
        if (sv[u] == 0.0f) continue;  // no interaction (or a zero term)
        const unsigned e = ev[u];
        const float4 pj = pos4[(int)(e & MFA_JMASK)];
        const float4 o = s_off[e >> MFA_JBITS];
        const float dx = (pi4.x - pj.x) + o.x;
        const float dy = (pi4.y - pj.y) + o.y;
        const float dz = (pi4.z - pj.z) + o.z;
        ax += (double)truncf(sv[u] * dx);
        ay += (double)truncf(sv[u] * dy);
        az += (double)truncf(sv[u] * dz);
      }
    }
  }
  if (active && sub == 0) {
    ax += hf.fpart[3 * (size_t)ss + 0];
    ay += hf.fpart[3 * (size_t)ss + 1];
    az += hf.fpart[3 * (size_t)ss + 2];
  }
  const bool bad = fix_bad(ax, ay, az);
  double fx = group_sum<TPP>(ax) * MFA_FIX_INV;
  double fy = group_sum<TPP>(ay) * MFA_FIX_INV;
  double fz = group_sum<TPP>(az) * MFA_FIX_INV;
  if (active && sub == 0 && !FUSE) {
    if (mask & TM_ACCUM) {
      fx = f[3 * (size_t)i + 0] + fx;
      fy = f[3 * (size_t)i + 1] + fy;
      fz = f[3 * (size_t)i + 2] + fz;
    }
    f[3 * (size_t)i + 0] = fx;
    f[3 * (size_t)i + 1] = fy;
    f[3 * (size_t)i + 2] = fz;
  }
  if (bad) atomicCAS(&flags[F_ERROR], 0, MFA_ERR_NAN);
  if (FUSE) fused_epilogue<TPP, true, false>(k, active, sub, i, il, ti, ss, fx, fy, fz, r, fu);
}

// ---------------------------------------------------------------------------
// Ewald sum in k space (ewald_k.f90:150-199) and the slab dipole correction
// (dipolar_correction.f90:20-33), FP64.  The k vectors of the half space
// (ewald_k.f90:104-124) are grouped into segments of at most EW_SEG
// consecutive kx at fixed (ky, kz).  Per charged particle j and step the
// phase factors factorise, exp(i k.r_j) = X_j(nx) Y_j(ny) Z_j(nz), and are
// tabulated once per axis (one sincos per entry); every (k, j) term is then a
// product of table entries, with no per-k reduction across lanes and no
// serial recurrence.  Charged particles only (q = 0 contributes nothing).
#define EW_SEG 8   // k vectors per segment
#define EW_CH 32   // charged particles per structure-factor block

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
#pragma clang fp contract(fast)
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Phase tables X[n][j] (n = nx + nxm), Y[n][j] (n = ny + nym), Z[n][j]
// (n = nz) of the charged particles: one thread per particle and axis (block
// row blockIdx.y = axis), one sincos of dk r, then the powers by complex
// products, e^{i n t} = e^{i (n-1) t} e^{i t} (n ulp of drift at most, for
// n <= nmax), and the negative n as conjugates; entries of consecutive
// particles in consecutive threads.  The x blocks also write the partial
// dipole moments sum q (z - lz_half).
__global__ __launch_bounds__(MFA_BLOCK) void k_ewald_tables(EwaldK e, const double* __restrict__ r,
                                                            const int* __restrict__ qidx,
                                                            const double* __restrict__ qval,
                                                            double2* __restrict__ tx,
                                                            double2* __restrict__ ty,
                                                            double2* __restrict__ tz,
                                                            double* __restrict__ mz_part) {
#pragma clang fp contract(fast)  // FP64 sums within 1e-12: fused multiply-adds
  __shared__ double sh[MFA_BLOCK / 64];
  const int idx = blockIdx.x * MFA_BLOCK + threadIdx.x;  // charged particle
  const int ax = blockIdx.y;                              // 0 x, 1 y, 2 z
  double mz = 0.0;
  if (idx < e.nq) {
    const int i = qidx[idx];
    const double x = r[3 * (size_t)i + ax];
    double s1, c1;
    sincos(e.dk[ax] * x, &s1, &c1);
    const double2 t1 = make_double2(c1, s1);
    double2 t = make_double2(1.0, 0.0);
    const size_t nq = (size_t)e.nq;
    if (ax == 2) {
      tz[idx] = t;
      for (int n = 1; n <= e.nzm; ++n) {
        t = cmul(t, t1);
        tz[(size_t)n * nq + idx] = t;
      }
    } else {
      double2* tab = ax == 0 ? tx : ty;
      const int m = ax == 0 ? e.nxm : e.nym;
      tab[(size_t)m * nq + idx] = t;
      for (int n = 1; n <= m; ++n) {
        t = cmul(t, t1);
        tab[(size_t)(m + n) * nq + idx] = t;
        tab[(size_t)(m - n) * nq + idx] = make_double2(t.x, -t.y);
      }
    }
    if (ax == 0) mz = qval[idx] * (r[3 * (size_t)i + 2] - e.lz_half);
  }
  if (ax != 0) return;
  for (int o = 32; o > 0; o >>= 1) mz += __shfl_down(mz, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mz;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < MFA_BLOCK / 64; ++w) m += sh[w];
    mz_part[blockIdx.x] = m;
  }
}

// Partial structure factor of one chunk of EW_CH charged particles: thread t
// of block (chunk, g) owns segment 256 g + t and accumulates
// sum_j q_j X_j(nx) Y_j(ny) Z_j(nz) over the chunk for its k vectors; the
// chunk's table rows are staged in LDS particle-major, each followed by
// EW_SEG zeros so that a segment always reads EW_SEG X entries (straight-line
// code: the ten reads of a particle are in flight together; sums past the
// segment's length are dropped).
__global__ __launch_bounds__(MFA_BLOCK) void k_ewald_rho(EwaldK e, const double* __restrict__ qval,
                                                         const double2* __restrict__ tx,
                                                         const double2* __restrict__ ty,
                                                         const double2* __restrict__ tz,
                                                         const int4* __restrict__ segs,
                                                         const int* __restrict__ seg_k0,
                                                         double2* __restrict__ rho_part) {
#pragma clang fp contract(fast)  // FP64 sums within 1e-12: fused multiply-adds
  extern __shared__ double2 ew_lds[];  // [EW_CH][nex + ney + nez + EW_SEG]
  const int row = e.nex + e.ney + e.nez + EW_SEG;
  const int j0 = blockIdx.x * EW_CH;
  const int nj = min(EW_CH, e.nq - j0);
  for (int t = threadIdx.x; t < nj * row; t += MFA_BLOCK) {
    const int j = t % nj, n = t / nj;  // coalesced over j
    double2 v;
    if (n >= e.nex + e.ney + e.nez) {
      v = make_double2(0.0, 0.0);
    } else if (n < e.nex) {
      v = tx[(size_t)n * e.nq + j0 + j];
    } else if (n < e.nex + e.ney) {
      v = ty[(size_t)(n - e.nex) * e.nq + j0 + j];
    } else {
      const double q = qval[j0 + j];
      const double2 zz = tz[(size_t)(n - e.nex - e.ney) * e.nq + j0 + j];
      v = make_double2(q * zz.x, q * zz.y);
    }
    ew_lds[j * row + n] = v;
  }
  __syncthreads();
  const int sg = blockIdx.y * MFA_BLOCK + threadIdx.x;
  if (sg >= e.nsegs) return;
  const int4 s = segs[sg];  // (ny, nz, nx_lo, length)
  const int yo = e.nex + s.x + e.nym, zo = e.nex + e.ney + s.y, xo = s.z + e.nxm;
  double2 acc[EW_SEG];
  for (int u = 0; u < EW_SEG; ++u) acc[u] = make_double2(0.0, 0.0);
  for (int j = 0; j < nj; ++j) {
    const double2* tj = ew_lds + j * row;
    const double2 yv = tj[yo], zv = tj[zo];
    double2 xv[EW_SEG];
#pragma unroll
    for (int u = 0; u < EW_SEG; ++u) xv[u] = tj[xo + u];
    const double2 a = cmul(yv, zv);
#pragma unroll
    for (int u = 0; u < EW_SEG; ++u) {
      acc[u].x = fma(xv[u].x, a.x, acc[u].x);
      acc[u].x = fma(-xv[u].y, a.y, acc[u].x);
      acc[u].y = fma(xv[u].x, a.y, acc[u].y);
      acc[u].y = fma(xv[u].y, a.x, acc[u].y);
    }
  }
  const int k0 = seg_k0[sg];
  for (int u = 0; u < EW_SEG; ++u)
    if (u < s.w) rho_part[(size_t)blockIdx.x * e.nk + k0 + u] = acc[u];
}

// rho(k) = sum of the chunk partials (fixed order), stored weighted by
// g_k = exp(-k^2/(4 alpha_c))/k^2; per-block sums of g_k |rho(k)|^2 for the
// energy; block 0 also sums the dipole moment.  A block of EW_SUMW waves
// covers 64 k vectors: wave w adds the partials c = w, w + EW_SUMW, ..., then
// wave 0 combines the wave sums in a fixed order.
#define EW_SUMW 16
__global__ __launch_bounds__(64 * EW_SUMW) void k_ewald_sum(EwaldK e,
                                                            const double2* __restrict__ rho_part,
                                                            const double4* __restrict__ kvec,
                                                            double2* __restrict__ grho,
                                                            double* __restrict__ uk_part,
                                                            const double* __restrict__ mz_part,
                                                            double* __restrict__ mz) {
#pragma clang fp contract(fast)  // FP64 sums within 1e-12: fused multiply-adds
  __shared__ double2 sh[EW_SUMW][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int kk = blockIdx.x * 64 + lane;
  double re = 0.0, im = 0.0;
  if (kk < e.nk) {
    int cb = w;
    for (; cb + 3 * EW_SUMW < e.n_chunks; cb += 4 * EW_SUMW) {  // four loads in flight
      double2 t[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) t[u] = rho_part[(size_t)(cb + u * EW_SUMW) * e.nk + kk];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        re += t[u].x;
        im += t[u].y;
      }
    }
    for (; cb < e.n_chunks; cb += EW_SUMW) {
      const double2 t = rho_part[(size_t)cb * e.nk + kk];
      re += t.x;
      im += t.y;
    }
  }
  sh[w][lane] = make_double2(re, im);
  __syncthreads();
  if (w == 0) {
    double u = 0.0;
    if (kk < e.nk) {
      re = 0.0;
      im = 0.0;
      for (int q = 0; q < EW_SUMW; ++q) {
        re += sh[q][lane].x;
        im += sh[q][lane].y;
      }
      const double g = kvec[kk].w;
      grho[kk] = make_double2(g * re, g * im);
      u = g * (re * re + im * im);
    }
    for (int o = 32; o > 0; o >>= 1) u += __shfl_down(u, o, 64);
    if (lane == 0) {
      uk_part[blockIdx.x] = u;
      if (blockIdx.x == 0) {
        double m = 0.0;
        for (int b = 0; b < e.n_mz; ++b) m += mz_part[b];
        *mz = m;
      }
    }
  }
}

// Sum over k of k g_k (cos(k.r) Im rho - sin(k.r) Re rho) for each charged
// particle (ewald_k.f90:177-192 without the prefactor).  A block holds 64
// particles (the lanes) and 4 waves; their X table columns and g_k rho(k) of
// the block's segments (group blockIdx.y) are staged in LDS.  The segments
// are dealt to the waves in turn, all lanes of a wave on the same segment;
// each wave loads the Y and Z entries of its next segment while it works on
// the current one.  The 4 wave sums are combined in a fixed order into one
// partial per block.
__global__ __launch_bounds__(MFA_BLOCK) void k_ewald_force(EwaldK e, const double2* __restrict__ tx,
                                                           const double2* __restrict__ ty,
                                                           const double2* __restrict__ tz,
                                                           const int4* __restrict__ segs,
                                                           const int* __restrict__ seg_k0,
                                                           const double2* __restrict__ grho,
                                                           double* __restrict__ fk_part) {
#pragma clang fp contract(fast)  // FP64 sums within 1e-12: fused multiply-adds
  extern __shared__ double2 ew_xs[];  // [nex][64], then [segs_per_kgroup][EW_SEG]
  __shared__ double sh[4][64][3];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int idx = blockIdx.x * 64 + lane;
  const bool live = idx < e.nq;
  const int jq = live ? idx : 0;
  double2* gs = ew_xs + e.nex * 64;
  const int g0 = blockIdx.y * e.segs_per_kgroup, g1 = min(e.nsegs, g0 + e.segs_per_kgroup);
  for (int n = w; n < e.nex; n += 4) ew_xs[n * 64 + lane] = tx[(size_t)n * e.nq + jq];
  for (int t = threadIdx.x; t < (g1 - g0) * EW_SEG; t += MFA_BLOCK) {
    const int sg = g0 + t / EW_SEG, u = t % EW_SEG;
    const int4 s = segs[sg];
    gs[t] = u < s.w ? grho[seg_k0[sg] + u] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  double fx = 0.0, fy = 0.0, fz = 0.0;
  int sg = g0 + w;
  double2 yv = make_double2(0.0, 0.0), zv = yv;
  int4 s = make_int4(0, 0, 0, 0);
  if (sg < g1) {
    s = segs[sg];
    yv = ty[(size_t)(s.x + e.nym) * e.nq + jq];
    zv = tz[(size_t)s.y * e.nq + jq];
  }
  while (sg < g1) {
    const int nsg = sg + 4;
    int4 ns = make_int4(0, 0, 0, 0);
    double2 nyv = make_double2(0.0, 0.0), nzv = nyv;
    if (nsg < g1) {  // the next segment's entries, loaded ahead
      ns = segs[nsg];
      nyv = ty[(size_t)(ns.x + e.nym) * e.nq + jq];
      nzv = tz[(size_t)ns.y * e.nq + jq];
    }
    const double2 b = cmul(yv, zv);
    const double2* xs = ew_xs + (s.z + e.nxm) * 64 + lane;
    const double2* g = gs + (sg - g0) * EW_SEG;
    double tsum = 0.0, kx = e.dk[0] * (double)s.z;
    for (int u = 0; u < s.w; ++u, kx += e.dk[0]) {  // s.w is uniform across the wave
      const double2 xe = cmul(xs[u * 64], b);
      const double t = xe.x * g[u].y - xe.y * g[u].x;
      fx += kx * t;
      tsum += t;
    }
    fy += (e.dk[1] * (double)s.x) * tsum;
    fz += (e.dk[2] * (double)s.y) * tsum;
    sg = nsg;
    s = ns;
    yv = nyv;
    zv = nzv;
  }
  sh[w][lane][0] = fx;
  sh[w][lane][1] = fy;
  sh[w][lane][2] = fz;
  __syncthreads();
  if (w == 0 && live) {
    double* o = fk_part + fk_index(idx, blockIdx.y, gridDim.y);
    for (int d = 0; d < 3; ++d)
      o[d] = (sh[0][lane][d] + sh[1][lane][d]) + (sh[2][lane][d] + sh[3][lane][d]);
  }
}

// The same sums with the k vectors taken row by row (the maximal runs of
// consecutive kx at fixed (ky, kz)): each lane owns EW_PPL charged particles
// (lane + 256 u of its block's 256 EW_PPL), their X table rows (all
// 2 NXM + 1 entries) held in registers; the block's rows of G = g_k rho(k)
// are staged in LDS as full rows of 2 NXM + 1 entries (zero where a row has
// no vector) and read by the whole wave at once.  Per k vector
// t = Im(conj(X Y Z) G) = cos(k.r) Im(g rho) - sin(k.r) Re(g rho); per row the
// sums factorise as
//   sum_n t_n       = Im(conj(Y Z) S),  S  = sum_n conj(X_n) G_n,
//   sum_n kx_n t_n  = dk_x Im(conj(Y Z) S'),  S' = sum_n n conj(X_n) G_n,
// eight FMA-class operations per k vector and particle.  The row's entries
// are one scalar bit mask; the unrolled sums test one bit per entry and skip
// the entries outside the row by a wave-uniform branch, with the entries in
// the row on the fall-through path.
// The launch is one-dimensional and XCD-aware: blocks b, b + 8, ... run on one
// XCD (its own L2), so each XCD gets a contiguous range of (particle block,
// row group) pairs, particle-block major, and its L2 holds the X, Y, Z table
// entries of its own particles only.  Row group g covers rows
// [g rows_per_group, (g + 1) rows_per_group).
#define EW_PPL 1
template <int NXM>
__global__ __launch_bounds__(MFA_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_ewald_force_rows(
    EwaldK e, const double2* __restrict__ tx, const double2* __restrict__ ty,
    const double2* __restrict__ tz, const int4* __restrict__ rows, const int* __restrict__ row_k0,
    int nrows, int rows_per_group, const double2* __restrict__ grho, double* __restrict__ fk_part) {
#pragma clang fp contract(fast)  // FP64 sums within 1e-12: fused multiply-adds
  constexpr int NEX = 2 * NXM + 1;
  constexpr int PB = MFA_BLOCK * EW_PPL;  // particles per block
  extern __shared__ double2 ew_g[];       // [rows_per_group][NEX]
  const int ngroups = (nrows + rows_per_group - 1) / rows_per_group;
  const int L = (int)(blockIdx.x & 7) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
  const int pblk = L / ngroups, grp = L - pblk * ngroups;
  if (pblk * PB >= e.nq) return;  // padding of the grid to a multiple of 8
  const int r0 = grp * rows_per_group, r1 = min(nrows, r0 + rows_per_group);
  for (int t = threadIdx.x; t < (r1 - r0) * NEX; t += MFA_BLOCK) {
    const int rl = t / NEX, nx = t % NEX - NXM;
    const int4 rw = rows[r0 + rl];
    ew_g[t] = (nx >= rw.z && nx <= rw.w) ? grho[row_k0[r0 + rl] + (nx - rw.z)] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  int idx[EW_PPL], jq[EW_PPL];
  double xr[EW_PPL][NEX], xi[EW_PPL][NEX];
  double fx[EW_PPL], fy[EW_PPL], fz[EW_PPL];
#pragma unroll
  for (int u = 0; u < EW_PPL; ++u) {
    idx[u] = pblk * PB + u * MFA_BLOCK + threadIdx.x;
    jq[u] = idx[u] < e.nq ? idx[u] : 0;
    fx[u] = fy[u] = fz[u] = 0.0;
#pragma unroll
    for (int n = 0; n < NEX; ++n) {
      const double2 v = tx[(size_t)n * e.nq + jq[u]];
      xr[u][n] = v.x;
      xi[u][n] = v.y;
    }
  }
  // (ny, nz, nx_lo, nx_hi) of a row: the same for the whole block
  auto row_of = [&](int r) {
    const int4 q = rows[r < r1 ? r : r0];
    return make_int4(__builtin_amdgcn_readfirstlane(q.x), __builtin_amdgcn_readfirstlane(q.y),
                     __builtin_amdgcn_readfirstlane(q.z), __builtin_amdgcn_readfirstlane(q.w));
  };
  int4 nrw = row_of(r0);
  double2 ny_[EW_PPL], nz_[EW_PPL];
#pragma unroll
  for (int u = 0; u < EW_PPL; ++u) {
    ny_[u] = ty[(size_t)(nrw.x + e.nym) * e.nq + jq[u]];
    nz_[u] = tz[(size_t)nrw.y * e.nq + jq[u]];
  }
  for (int r = r0; r < r1; ++r) {
    const int4 rw = nrw;
    double2 y[EW_PPL], z[EW_PPL];
    nrw = row_of(r + 1);  // the next row's table entries, fetched during this one
#pragma unroll
    for (int u = 0; u < EW_PPL; ++u) {
      y[u] = ny_[u];
      z[u] = nz_[u];
      ny_[u] = ty[(size_t)(nrw.x + e.nym) * e.nq + jq[u]];
      nz_[u] = tz[(size_t)nrw.y * e.nq + jq[u]];
    }
    // the row's vectors only (nx_lo..nx_hi, the same for the whole wave)
    const double2* g = ew_g + (r - r0) * NEX;
    double sr[EW_PPL], si[EW_PPL], pr[EW_PPL], pi[EW_PPL];
#pragma unroll
    for (int u = 0; u < EW_PPL; ++u) sr[u] = si[u] = pr[u] = pi[u] = 0.0;
    // the row's entries as one scalar bit mask: one bit test per entry
    const unsigned in_row = ((2u << (rw.w - rw.z)) - 1u) << (rw.z + NXM);
#pragma unroll
    for (int n = 0; n < NEX; ++n) {
      if (__builtin_expect((in_row & (1u << n)) == 0u, 0)) continue;  // entries in the row fall through
      const double2 gv = g[n];
#pragma unroll
      for (int u = 0; u < EW_PPL; ++u) {
        const double a = xr[u][n] * gv.x + xi[u][n] * gv.y;
        const double b = xr[u][n] * gv.y - xi[u][n] * gv.x;
        sr[u] += a;
        si[u] += b;
        if (n != NXM) {
          pr[u] += (double)(n - NXM) * a;
          pi[u] += (double)(n - NXM) * b;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < EW_PPL; ++u) {
      const double br = y[u].x * z[u].x - y[u].y * z[u].y, bi = y[u].x * z[u].y + y[u].y * z[u].x;
      const double tsum = br * si[u] - bi * sr[u];
      fx[u] += e.dk[0] * (br * pi[u] - bi * pr[u]);
      fy[u] += (e.dk[1] * (double)rw.x) * tsum;
      fz[u] += (e.dk[2] * (double)rw.y) * tsum;
    }
  }
#pragma unroll
  for (int u = 0; u < EW_PPL; ++u)
    if (idx[u] < e.nq) {
      double* o = fk_part + fk_index(idx[u], grp, ngroups);
      o[0] = fx[u];
      o[1] = fy[u];
      o[2] = fz[u];
    }
}
#define EW_ROWS_MAX_NXM 12  // register rows up to 25 X entries per particle
// force partials (one per block row and charged particle) the buffer holds
#define EW_FK_CAP(n) ((size_t)2048 * 64 + (size_t)64 * (size_t)(n))

// Dipole moment of the charged particles, one block (the per-subroutine
// entry point; the step loop takes it from k_ewald_rho/k_ewald_sum).
__global__ __launch_bounds__(MFA_BLOCK) void k_ewald_mz(EwaldK e, const double* __restrict__ r,
                                                        const int* __restrict__ qidx,
                                                        const double* __restrict__ qval,
                                                        double* __restrict__ mz) {
  __shared__ double sh[MFA_BLOCK];
  double m = 0.0;
  for (int idx = threadIdx.x; idx < e.nq; idx += MFA_BLOCK)
    m += qval[idx] * (r[3 * (size_t)qidx[idx] + 2] - e.lz_half);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int o = MFA_BLOCK / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *mz = sh[0];
}

// Adds -8 pi q f inv_V (ewald_k.f90:195-196) and the slab dipole force
// -4 pi q M_z inv_V (dipolar_correction.f90:29) to F.
// One thread per particle and component; the k-group partials are summed in
// group order, eight loads in flight.
__global__ void k_ewald_apply(EwaldK e, const int* __restrict__ qidx,
                              const double* __restrict__ qval, const double* __restrict__ fk_part,
                              int n_kgroups, const double* __restrict__ mz, double inv_V,
                              int add_k, int add_dip, double* __restrict__ f) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * e.nq) return;
  const int idx = t / 3, d = t - 3 * idx;
  const int i = qidx[idx];
  const double sk = add_k ? fk_sum(fk_part, idx, n_kgroups, d) : 0.0;
  f[3 * (size_t)i + d] = kspace_add(f[3 * (size_t)i + d], sk, qval[idx], inv_V, add_k, add_dip, d, mz);
}

// u_cou_k = 4 pi inv_V sum g_k |rho_k|^2 (ewald_k.f90:160-170) and the dipole
// energy 2 pi M_z^2 inv_V (dipolar_correction.f90:32).
__global__ void k_ewald_energy(const double* __restrict__ uk_part, int nparts,
                               const double* __restrict__ mz, double inv_V,
                               double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double pi_f = (double)3.1415926536f;
  double u = 0.0;
  for (int b = 0; b < nparts; ++b) u += uk_part[b];
  out[0] = u * 4.0 * pi_f * inv_V;
  out[1] = 2.0 * pi_f * (*mz) * (*mz) * inv_V;
}

// ---------------------------------------------------------------------------
// DPD_VV (md_main.f90:149-155, new_dpd_fd.f90:18-177, DPD_WEIGHT 0): after
// verlet_velocities the dissipative force is computed again with the new
// velocities, for list pairs inside range_2 with neither bead a head (type
// 1), images in x and y only (new_dpd_fd.f90:38-44), w_d =
// (1 - sqrt(r_2/range_2))^2; F_D,ij = -gamma w_d (dv.e) e.  Each particle
// gathers its own terms from its full list row (sorted order, TPP lanes as
// in the force pass); F_D,ji = -F_D,ij bit for bit.  FP64 in both forms.
template <int TPP>
__global__ __launch_bounds__(MFA_BLOCK) void k_dpd_fd(const KParams* __restrict__ kb,
                                                      const int* __restrict__ rmap,
                                                      const double* __restrict__ r,
                                                      const double* __restrict__ v,
                                                      const int* __restrict__ type,
                                                      const int* __restrict__ perm,
                                                      const unsigned* __restrict__ nl,
                                                      const int* __restrict__ nn, int cap,
                                                      double* __restrict__ fd_new) {
  MFA_REPLICA(blockIdx.x * (MFA_BLOCK / TPP));
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int sp = gtid / TPP, sub = gtid % TPP;
  const bool active = sp - k.p0 < k.n;
  const int ss = active ? sp : k.p0;
  const int i = perm[ss];
  const int ti = type[i] - 1;
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (active && ti != 0) {
    const double xi = r[3 * (size_t)i + 0], yi = r[3 * (size_t)i + 1], zi = r[3 * (size_t)i + 2];
    const double vxi = v[3 * (size_t)i + 0], vyi = v[3 * (size_t)i + 1], vzi = v[3 * (size_t)i + 2];
    const int wn = nn[ss], cnt = wn & 0xffff, nup = wn >> 16;
    const unsigned* row = nl + (size_t)ss * cap;
    for (int q = sub; q < cnt; q += TPP) {
      const int j = perm[row[list_slot(q, nup, cnt, cap)] & MFA_JMASK];
      const int tj = type[j] - 1;
      if (tj == 0) continue;
      double dx = pbc(xi - r[3 * (size_t)j + 0], k.L[0], k.invL[0]);
      double dy = pbc(yi - r[3 * (size_t)j + 1], k.L[1], k.invL[1]);
      const double dz = zi - r[3 * (size_t)j + 2];
      const double r2 = (dx * dx + dy * dy) + dz * dz;
      if (!(r2 < k.range_2[ti][tj])) continue;
      const double wr = 1.0 - sqrt(r2 / k.range_2[ti][tj]);
      const double wd = wr * wr;
      const double rn = sqrt(r2);
      const double ex = dx / rn, ey = dy / rn, ez = dz / rn;
      const double rd = ((vxi - v[3 * (size_t)j + 0]) * ex + (vyi - v[3 * (size_t)j + 1]) * ey) +
                        (vzi - v[3 * (size_t)j + 2]) * ez;
      const double c = (-1.0 * k.friction1) * wd * rd;
      ax += c * ex;
      ay += c * ey;
      az += c * ez;
    }
  }
  ax = group_sum<TPP>(ax);
  ay = group_sum<TPP>(ay);
  az = group_sum<TPP>(az);
  if (active && sub == 0) {
    fd_new[3 * (size_t)i + 0] = ax;
    fd_new[3 * (size_t)i + 1] = ay;
    fd_new[3 * (size_t)i + 2] = az;
  }
}

// The rest of new_dpd_fd (:18-22,103-124,148-176): F = (F - F_D,old) + F_D,new,
// F_D,old <- F_D,new, v += (dt/2) F/m, a = F/m; the view's velocity; the
// kinetic energy partials; with inc_step the step counter moves on (this
// kernel then ends the step, as the fused integrate kernel does otherwise).
__global__ __launch_bounds__(MFA_BLOCK) void k_dpd_vv_apply(
    const KParams* __restrict__ kb, const int* __restrict__ rmap, double* __restrict__ f,
    double* __restrict__ fd, const double* __restrict__ fd_new, double* __restrict__ v,
    double* __restrict__ a, const int* __restrict__ type, const int* __restrict__ sidx,
    float4* __restrict__ vel4, double* __restrict__ part, int energy,
    long long* __restrict__ d_step, int inc_step) {
  MFA_REPLICA(blockIdx.x * blockDim.x);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (inc_step && i == 0) *d_step += 1;
  double vals[MFA_NTERMS];
  for (int t = 0; t < MFA_NTERMS; ++t) vals[t] = 0.0;
  if (i - k.p0 < k.n) {
    const int ti = type[i] - 1;
    const double im = k.inv_mass_type[ti];
    double vi[3];
    double tk = 0.0;
    for (int d = 0; d < 3; ++d) {
      const size_t q = 3 * (size_t)i + d;
      const double fn = fd_new[q];
      const double fi = (f[q] - fd[q]) + fn;
      fd[q] = fn;
      f[q] = fi;
      vi[d] = v[q] + (k.half_dt * fi) * im;
      v[q] = vi[d];
      a[q] = fi * im;
      tk += k.mass_type[ti] * (vi[d] * vi[d]);
    }
    vel4[sidx[i]] = make_float4((float)vi[0], (float)vi[1], (float)vi[2], 0.0f);
    vals[T_TKIN] = tk;
  }
  if (energy) block_partials(vals, part);
}

// ---------------------------------------------------------------------------
// Host side

// The skin-dependent kernel constants of a replica: the list criterion, the
// skin test and the FP32 decision margins (which bound coordinates by the
// list radius and the skin).
static void kparams_skin(KParams& k, const mfa_params& p) {
  for (int t = 0; t < MFA_NTYPE; ++t)
    for (int u = 0; u < MFA_NTYPE; ++u) k.list_2[t][u] = p.list_2[t][u];
  k.skin_half_2 = p.skin_half_2;
  k.skin_2 = p.skin_2;
  {
    // FP32 separation from FP32 positions: the two coordinates, the raw
    // difference (|d| <= L before the image shift) and the FP32 box edge each
    // carry at most ulp(L)/2 <= 2 L 2^-24, the shifted result |d| <= r_list
    // one more rounding.  Four times the resulting bound on r^2, plus the
    // rounding of the sum of squares and of list_2 to FP32.
    double lmax = 0.0, l2max = 0.0;
    for (int d = 0; d < 3; ++d) lmax = p.boundary[d] > lmax ? p.boundary[d] : lmax;
    for (int t = 0; t < MFA_NTYPE; ++t)
      for (int u = 0; u < MFA_NTYPE; ++u) l2max = p.list_2[t][u] > l2max ? p.list_2[t][u] : l2max;
    const double u = ldexp(1.0, -24);
    const double dcoord = 8.0 * lmax * u + 2.0 * sqrt(l2max) * u;
    k.pf_margin = (float)(4.0 * (3.0 * 2.0 * sqrt(l2max) * dcoord + 3.0 * dcoord * dcoord) +
                          8.0 * l2max * u);
    // Force pass: coordinates relative to the cell origin at the rebuild stay
    // within [-skin/2, r_bin + skin/2] (a particle that moves farther triggers
    // a rebuild); each of the two coordinates, their difference, the FP32 cell
    // offset and the final sum carry one rounding.
    double bmax = 0.0, rbmax = 0.0;
    for (int d = 0; d < 3; ++d) {
      bmax = p.r_bin[d] + p.skin > bmax ? p.r_bin[d] + p.skin : bmax;
      rbmax = p.r_bin[d] > rbmax ? p.r_bin[d] : rbmax;
    }
    const double dsep = (6.0 * bmax + 2.0 * rbmax) * u;
    double c2max = p.r_cut_dpd_2 > p.rc_coul_2 ? p.r_cut_dpd_2 : p.rc_coul_2;
    for (int t = 0; t < MFA_NTYPE; ++t)
      for (int q = 0; q < MFA_NTYPE; ++q) c2max = p.range_2[t][q] > c2max ? p.range_2[t][q] : c2max;
    k.ff_margin = (float)(4.0 * (2.0 * sqrt(3.0 * c2max) * dsep + 3.0 * dsep * dsep +
                                 6.0 * c2max * u));
    k.lf_margin = (float)(4.0 * (2.0 * sqrt(3.0 * l2max) * dsep + 3.0 * dsep * dsep +
                                 6.0 * l2max * u));
  }
}

// List capacity of a replica: the expected full-list count at its mean
// density, with room for fluctuations.
static int list_cap_of(const mfa_params& p) {
  const double dens = (double)p.n_part / (p.boundary[0] * p.boundary[1] * p.boundary[2]);
  double rl2 = 0.0;
  for (int t = 0; t < MFA_NTYPE; ++t)
    for (int u = 0; u < MFA_NTYPE; ++u) rl2 = p.list_2[t][u] > rl2 ? p.list_2[t][u] : rl2;
  const double rl = sqrt(rl2);
  const double expect = dens * 4.18879020478639 * rl * rl * rl;
  return (((int)(2.0 * expect + 48.0)) + 31) / 32 * 32;
}

static KParams make_kparams(const mfa_params& p) {
  KParams k;
  memset(&k, 0, sizeof k);
  for (int d = 0; d < 3; ++d) {
    k.L[d] = p.boundary[d];
    k.invL[d] = p.inv_boundary[d];
    k.inv_r_bin[d] = p.inv_r_bin[d];
    k.r_bin[d] = p.r_bin[d];
    k.nb[d] = p.n_bin[d];
    k.va[d] = p.va_spring_twall[d];
  }
  k.dt = p.dt;
  k.dt_2 = p.dt_2;
  k.half_dt = 0.5 * p.dt;      // 0.5*dt   (verlet_positions.f90:169)
  k.half_dt_2 = 0.5 * p.dt_2;  // 0.5*dt_2 (verlet_positions.f90:154)
  for (int t = 0; t < MFA_NTYPE; ++t) {
    k.inv_mass_type[t] = p.inv_mass_type[t];
    k.mass_type[t] = p.mass_type[t];
    k.sigma_wall[t] = p.sigma_wall[t];
    k.a_wall[t] = p.a_wall[t];
    for (int u = 0; u < MFA_NTYPE; ++u) {
      k.epsil[t][u] = p.epsil[t][u];
      k.sigma_2[t][u] = p.sigma_2[t][u];
      k.range_2[t][u] = p.range_2[t][u];
      k.inv_range_2[t][u] = p.inv_range_2[t][u];
      k.inv_range[t][u] = sqrt(p.inv_range_2[t][u]);
      k.e_shift[t][u] = p.e_shift[t][u];
    }
  }
  k.r_cut_dpd_2 = p.r_cut_dpd_2;
  k.inv_r_cut_dpd = p.r_cut_dpd_2 > 0.0 ? 1.0 / sqrt(p.r_cut_dpd_2) : 0.0;
  k.friction1 = p.friction1;
  k.sig = p.sig;
  k.k_chain = p.k_chain;
  k.inv_r_chain_2 = p.inv_r_chain_2;
  k.a_w = p.a_w;
  k.z_space_wall = p.z_space_wall;
  k.fso_r2_init = p.fso_r2_init;
  k.fso_time = p.fso_time;
  k.coulomb = p.coulomb;
  k.bending = p.bending;
  k.k_bend = p.k_bend;
  k.alpha_eq = p.alpha_eq;
  for (int t = 0; t < MFA_NTYPE; ++t) k.q_type[t] = p.charge[t];
  k.alpha_c = p.alpha_c;
  k.sqrt_alpha = p.sqrt_alpha;
  k.inv_sqrt_pi = p.inv_sqrt_pi_real;
  k.rc_coul_2 = p.coulomb ? p.rc_coul_2 : 0.0;
  kparams_skin(k, p);
  k.n_cells = p.n_bin[0] * p.n_bin[1] * p.n_bin[2];
  k.n = p.n_part;
  k.n_mon = p.n_mon > 0 ? p.n_mon : 1;
  k.n_chain = p.n_chain;
  k.n_brush = p.n_mon * p.n_chain;
  k.part_init_d = p.part_init_d;
  k.n_mon_d = p.n_mon_d > 0 ? p.n_mon_d : 1;
  k.part_init_e = p.part_init_e;
  k.n_mon_e = p.n_mon_e;
  k.symmetry = p.symmetry;
  // Heads pinned for brush systems in a channel (verlet_velocities.f90:113-134).
  k.pinned_heads = (p.system == 3 && p.symmetry == 0) ? 1 : 0;
  k.dpd_on = p.thermostat == 0 ? 1 : 0;
  k.lgv = p.thermostat == 1 ? 1 : 0;
  k.dpd_cut_off_on = p.dpd_cut_off_on;
  // DPD_CUT_OFF excludes heads for brush systems in a channel (fluid_fluid.f90:179-185).
  k.dpd_head_excl = (p.system == 3 && p.symmetry == 0) ? 1 : 0;
  k.seed = (uint32_t)p.seed;
  k.dpd_vv = p.dpd_vv;
  return k;
}

static int grid_of(int n) { return (n + MFA_BLOCK - 1) / MFA_BLOCK; }
static int launch_ewald(mfa_ctx* c, int add_dip);



// Candidates the list build stages in LDS at a time: all 27 cells' worth of
// a full cell capacity, at most LB_PIECE.
static int list_piece(const mfa_ctx* c) {
  const int all = 27 * c->cell_cap;
  return all < LB_PIECE ? all : LB_PIECE;
}

// Dynamic LDS of the list build: the cell's particles, the staged candidates
// (float4 + entry each) and the queues of candidate indices (per wave and
// group, LB_RING).
static size_t list_lds_bytes(const mfa_ctx* c) {
  return (size_t)c->cell_cap * sizeof(float4) +
         (size_t)list_piece(c) * (sizeof(float4) + sizeof(unsigned)) +
         (size_t)LB_WAVES * 2 * LB_RING * sizeof(int);
}

// A rebuild: cells of every particle, their sorted ranges, the spatial order
// and FP32 view, then the lists.  Each kernel returns at once unless the
// replica's rebuild flag is set.
// With ev (four events), the kernels are bracketed by them for timing.
// inc_step: the step counter moves on at the start of the rebuild (the steps
// whose force pass also integrates, which therefore cannot).
static int launch_rebuild(mfa_ctx* c, hipEvent_t* ev = nullptr, int inc_step = 0) {
  const int g = grid_of(c->n);
  if (ev) HIPCHK(c, hipEventRecord(ev[0], c->stream));
  hipLaunchKernelGGL(k_cell_fill, dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->d_rmap, c->r,
                     c->perm, c->cell_of, c->cell_count, c->cell_slots, c->cell_start, c->flags,
                     c->fill_cnt, c->d_step, inc_step);
  if (ev) HIPCHK(c, hipEventRecord(ev[1], c->stream));
  const int cb = strided_grid((c->n_cells + 3) / 4, SORT_GRID);  // groups of four cells (waves)
  hipLaunchKernelGGL(k_cell_sort, dim3(cb), dim3(MFA_BLOCK), (size_t)4 * c->cell_cap * sizeof(int),
                     c->stream, c->d_kp, c->d_cell_rep, c->n_cells, c->n_rep, c->r, c->v, c->type, c->mic,
                     c->cell_count, c->cell_slots, c->cell_start, c->perm, c->sidx, c->r_ref,
                     c->flags, c->pos4, c->vel4);
  if (ev) HIPCHK(c, hipEventRecord(ev[2], c->stream));
  const int lg = strided_grid(c->n_cells, LB_GRID);
  if (c->list_kind == LK_HALF)
    hipLaunchKernelGGL(k_build_list<LK_HALF>, dim3(lg), dim3(MFA_BLOCK), list_lds_bytes(c), c->stream,
                       c->n_cells, c->n_rep, c->d_kp, c->d_cell_rep, c->r, c->cell_count, c->cell_start, c->perm,
                       c->pos4, c->nl, c->nn, c->cap, c->flags, c->lb_octants ? 1 : 0, list_piece(c));
  else if (c->list_kind == LK_NEAR)
    hipLaunchKernelGGL(k_build_list<LK_NEAR>, dim3(lg), dim3(MFA_BLOCK), list_lds_bytes(c), c->stream,
                       c->n_cells, c->n_rep, c->d_kp, c->d_cell_rep, c->r, c->cell_count, c->cell_start, c->perm,
                       c->pos4, c->nl, c->nn, c->cap, c->flags, c->lb_octants ? 1 : 0, list_piece(c));
  else
    hipLaunchKernelGGL(k_build_list<LK_PLAIN>, dim3(lg), dim3(MFA_BLOCK), list_lds_bytes(c), c->stream,
                       c->n_cells, c->n_rep, c->d_kp, c->d_cell_rep, c->r, c->cell_count, c->cell_start, c->perm,
                       c->pos4, c->nl, c->nn, c->cap, c->flags, c->lb_octants ? 1 : 0, list_piece(c));
  if (ev) HIPCHK(c, hipEventRecord(ev[3], c->stream));
  HIPCHK(c, hipGetLastError());
  return MFA_OK;
}

// The half-list form applies to the mixed pair passes without energies, at
// 4 or more threads per particle, once its buffers exist for this capacity.
static bool half_ready(const mfa_ctx* c, int mask) {
  return c->list_kind == LK_HALF && c->precision == MFA_MIXED && c->tpp >= 4 && (mask & TM_PAIR) &&
         !(mask & TM_ENERGY) && c->pair_s && c->half_cap == c->cap;
}

static bool ewald_active(const mfa_ctx* c);

// What the force pass of a step that also integrates reads and writes
// besides its inputs: the other position and view buffers, v and mic in
// place (or, under mfa_time_force_step, the timing copies), the skin flags,
// the cell counts, and for charged systems the k-space force partials.
static FuseArgs fuse_args(const mfa_ctx* c) {
  FuseArgs fu;
  memset(&fu, 0, sizeof fu);
  fu.r_out = c->r_alt;
  fu.v_in = c->v;
  fu.mic_in = c->mic;
  fu.v_out = c->probe_v ? c->probe_v : c->v_alt;
  fu.mic_out = c->probe_mic ? c->probe_mic : c->mic;
  fu.rebuild_flags = c->probe_flags ? c->probe_flags : c->flags;
  fu.r_ref = c->r_ref;
  fu.cell_of = c->cell_of;
  fu.pos4_out = c->pos4_alt;
  fu.vel4_out = c->vel4_alt;
  fu.cell_count = c->probe_count ? c->probe_count : c->cell_count;
  fu.nc_tot = c->n_cells;
  if (ewald_active(c)) {
    fu.qinv = c->d_qinv;
    fu.qval = c->d_qval;
    fu.fk_part = c->d_fk_part;
    fu.mz = c->d_mz;
    fu.inv_V = c->ew.inv_V;
    fu.nq = c->ew.e.nq;
    fu.n_kgroups = c->ew.n_kgroups;
    fu.add_dip = c->p.symmetry == 0 ? 1 : 0;  // md_main.f90:128-132
  }
  return fu;
}

template <int TPP, int XT, int Q>
static void launch_force_txq(mfa_ctx* c, int mask) {
  const int g = grid_of(c->n * TPP);
  if (c->precision == MFA_MIXED) {
    FuseArgs fu;
    memset(&fu, 0, sizeof fu);
    HalfArgs hf = {c->pair_s, c->mir, c->fpart};
    if constexpr (TPP >= 4) {
      if (half_ready(c, mask)) {
        // first pass: upper entries, bonds, walls; second: lower entries,
        // then F (or, in a step, the integration)
        hipLaunchKernelGGL((k_force_mixed<TPP, 3, XT, Q>), dim3(g), dim3(MFA_BLOCK), 0, c->stream,
                           c->d_kp, c->d_rmap, c->r, c->pos4, c->vel4, c->perm, c->nl, c->nn, c->cap,
                           c->f, c->d_step, mask, c->part, c->flags, fu, hf);
        if (mask & TM_FUSE) {
          fu = fuse_args(c);
          hipLaunchKernelGGL((k_force_lower<TPP, 2>), dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp,
                             c->d_rmap, c->r, c->pos4, c->perm, c->nl, c->nn, c->cap, c->f, mask,
                             c->flags, fu, hf);
        } else {
          hipLaunchKernelGGL((k_force_lower<TPP, 0>), dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp,
                             c->d_rmap, c->r, c->pos4, c->perm, c->nl, c->nn, c->cap, c->f, mask,
                             c->flags, fu, hf);
        }
        return;
      }
    }
    if (mask & TM_FUSE) {
      fu = fuse_args(c);
      hipLaunchKernelGGL((k_force_mixed<TPP, 2, XT, Q>), dim3(g), dim3(MFA_BLOCK), 0, c->stream,
                         c->d_kp, c->d_rmap, c->r, c->pos4, c->vel4, c->perm, c->nl, c->nn, c->cap, c->f,
                         c->d_step, mask, c->part, c->flags, fu, hf);
    } else if (mask & TM_ENERGY) {
      hipLaunchKernelGGL((k_force_mixed<TPP, 1, XT, Q>), dim3(g), dim3(MFA_BLOCK), 0, c->stream,
                         c->d_kp, c->d_rmap, c->r, c->pos4, c->vel4, c->perm, c->nl, c->nn, c->cap, c->f,
                         c->d_step, mask, c->part, c->flags, fu, hf);
    } else {
      hipLaunchKernelGGL((k_force_mixed<TPP, 0, XT, Q>), dim3(g), dim3(MFA_BLOCK), 0, c->stream,
                         c->d_kp, c->d_rmap, c->r, c->pos4, c->vel4, c->perm, c->nl, c->nn, c->cap, c->f,
                         c->d_step, mask, c->part, c->flags, fu, hf);
    }
  } else {
    FuseArgs fu;
    memset(&fu, 0, sizeof fu);
    if (mask & TM_FUSE) {
      fu = fuse_args(c);
      hipLaunchKernelGGL((k_force64<TPP, XT & 3, false, true>), dim3(g), dim3(MFA_BLOCK), 0, c->stream,
                         c->d_kp, c->d_rmap, c->r, c->v, c->type, c->perm, c->nl, c->nn, c->cap, c->f,
                         c->d_step, mask, c->part, c->flags, fu);
    } else if (mask & TM_ENERGY) {
      hipLaunchKernelGGL((k_force64<TPP, XT & 3, true, false>), dim3(g), dim3(MFA_BLOCK), 0, c->stream,
                         c->d_kp, c->d_rmap, c->r, c->v, c->type, c->perm, c->nl, c->nn, c->cap, c->f,
                         c->d_step, mask, c->part, c->flags, fu);
    } else {
      hipLaunchKernelGGL((k_force64<TPP, XT & 3, false, false>), dim3(g), dim3(MFA_BLOCK), 0, c->stream,
                         c->d_kp, c->d_rmap, c->r, c->v, c->type, c->perm, c->nl, c->nn, c->cap, c->f,
                         c->d_step, mask, c->part, c->flags, fu);
    }
  }
}

// The mixed pass's pair queue modes (compiled for DPD_CUT_OFF systems and
// for systems with no extra terms, the cases where many list candidates do
// not interact) when the context has one on.
template <int TPP, int XT>
static void launch_force_tx(mfa_ctx* c, int mask) {
  if constexpr ((XT & 4) != 0 || XT == 0) {
    if (c->pair_queue == 1 && c->precision == MFA_MIXED) {
      launch_force_txq<TPP, XT, 1>(c, mask);
      return;
    }
    if (c->pair_queue == 2 && c->precision == MFA_MIXED) {
      launch_force_txq<TPP, XT, 2>(c, mask);
      return;
    }
  }
  launch_force_txq<TPP, XT, 0>(c, mask);
}

// XT bit 1: Coulomb real space and bending; bit 2: the Langevin force; bit 4
// (mixed form): DPD_CUT_OFF.  The terms a system does not use are compiled
// out of its force kernels (they would otherwise cost registers, hence
// occupancy, and decision logic in every launch).
template <int TPP>
static void launch_force_t(mfa_ctx* c, int mask) {
  switch ((c->extra_terms ? 1 : 0) | (c->lgv_terms ? 2 : 0) | (c->p.dpd_cut_off_on ? 4 : 0)) {
    case 0: launch_force_tx<TPP, 0>(c, mask); break;
    case 1: launch_force_tx<TPP, 1>(c, mask); break;
    case 2: launch_force_tx<TPP, 2>(c, mask); break;
    case 3: launch_force_tx<TPP, 3>(c, mask); break;
    case 4: launch_force_tx<TPP, 4>(c, mask); break;
    case 5: launch_force_tx<TPP, 5>(c, mask); break;
    case 6: launch_force_tx<TPP, 6>(c, mask); break;
    default: launch_force_tx<TPP, 7>(c, mask); break;
  }
}

static int force_blocks(mfa_ctx* c) { return grid_of(c->n * c->tpp); }

static int timed_begin(mfa_ctx* c, std::vector<hipEvent_t>& pool, size_t& used) {
  while (pool.size() < used + 2) {
    hipEvent_t e;
    HIPCHK(c, hipEventCreate(&e));
    pool.push_back(e);
  }
  HIPCHK(c, hipEventRecord(pool[used], c->stream));
  return MFA_OK;
}
static int timed_end(mfa_ctx* c, std::vector<hipEvent_t>& pool, size_t& used) {
  HIPCHK(c, hipEventRecord(pool[used + 1], c->stream));
  used += 2;
  return MFA_OK;
}
// Adds the elapsed times of the recorded pairs to (ms, count).
static int timed_flush(mfa_ctx* c, std::vector<hipEvent_t>& pool, size_t& used, double& ms,
                       int64_t& count) {
  if (used == 0) return MFA_OK;
  HIPCHK(c, hipEventSynchronize(pool[used - 1]));
  for (size_t q = 0; q < used; q += 2) {
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, pool[q], pool[q + 1]));
    ms += t;
    count += 1;
  }
  used = 0;
  return MFA_OK;
}

static int launch_force(mfa_ctx* c, int mask) {
  const bool timed = c->timing && !c->capturing;
  if (timed) {
    if (c->ev_force_used + 2 > 4096) {
      int rc = timed_flush(c, c->ev_force, c->ev_force_used, c->force_ms, c->force_n);
      if (rc) return rc;
    }
    int rc = timed_begin(c, c->ev_force, c->ev_force_used);
    if (rc) return rc;
  }
  switch (c->tpp) {
    case 1: launch_force_t<1>(c, mask); break;
    case 2: launch_force_t<2>(c, mask); break;
    case 4: launch_force_t<4>(c, mask); break;
    case 8: launch_force_t<8>(c, mask); break;
    case 16: launch_force_t<16>(c, mask); break;
    default: return set_err(c, MFA_ERR_ARG, "threads per particle %d unsupported", c->tpp);
  }
  HIPCHK(c, hipGetLastError());
  if (timed) return timed_end(c, c->ev_force, c->ev_force_used);
  return MFA_OK;
}

static int launch_fused(mfa_ctx* c, int energy) {
  const bool timed = c->timing && !c->capturing;
  int rc;
  if (timed) {
    if (c->ev_integ_used + 2 > 4096) {
      rc = timed_flush(c, c->ev_integ, c->ev_integ_used, c->integ_ms, c->integ_n);
      if (rc) return rc;
    }
    rc = timed_begin(c, c->ev_integ, c->ev_integ_used);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_fused_integrate, dim3(grid_of(c->n)), dim3(MFA_BLOCK), 0, c->stream,
                     c->d_kp, c->d_rmap, c->f, c->v, c->r, c->mic, c->r_ref, c->type, c->flags,
                     c->part, energy, c->n_cells, c->cell_count, c->d_step, c->cell_of, c->sidx,
                     c->pos4, c->vel4);
  HIPCHK(c, hipGetLastError());
  if (timed) return timed_end(c, c->ev_integ, c->ev_integ_used);
  return MFA_OK;
}

static int set_step(mfa_ctx* c, int64_t step) {
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, c->stream, c->d_step, (long long)step);
  HIPCHK(c, hipGetLastError());
  return MFA_OK;
}

// Sums of the block partials [b0, b0 + nblocks) (one replica's blocks).
static int reduce_terms_range(mfa_ctx* c, int b0, int nblocks, double* out) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(MFA_NTERMS), dim3(MFA_BLOCK), 0, c->stream, c->part,
                     b0, nblocks, c->red);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->h_red, c->red, MFA_NTERMS * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int t = 0; t < MFA_NTERMS; ++t) out[t] = c->h_red[t];
  return MFA_OK;
}

// Blocks of the force pass (TPP lanes per particle) that cover replica q.
static void force_block_range(const mfa_ctx* c, int q, int* b0, int* nb) {
  const int per = MFA_BLOCK / c->tpp;  // particles per block; divides 256
  *b0 = c->ks[(size_t)q].p0 / per;
  *nb = (c->ks[(size_t)q].n + per - 1) / per;
}
static void particle_block_range(const mfa_ctx* c, int q, int* b0, int* nb) {
  *b0 = c->ks[(size_t)q].p0 / MFA_BLOCK;
  *nb = grid_of(c->ks[(size_t)q].n);
}

// Replica 0's sums of a force pass / an integrate pass.
static int reduce_terms(mfa_ctx* c, int nblocks, double* out) {
  (void)nblocks;
  int b0, nb;
  force_block_range(c, 0, &b0, &nb);
  return reduce_terms_range(c, b0, nb, out);
}
static int reduce_terms_particles(mfa_ctx* c, double* out) {
  int b0, nb;
  particle_block_range(c, 0, &b0, &nb);
  return reduce_terms_range(c, b0, nb, out);
}

static int check_flags(mfa_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->h_flags, c->flags, (size_t)c->n_rep * F_NFLAGS * sizeof(int32_t),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int q = 0; q < c->n_rep; ++q) {
    const int* h = c->h_flags + (size_t)q * F_NFLAGS;
    const int e = h[F_ERROR];
    if (e == 0) continue;
    char who[32] = "";
    if (c->n_rep > 1) snprintf(who, sizeof who, "replica %d: ", q);
    if (e == MFA_ERR_LIST_OVERFLOW)
      return set_err(c, e, "%sneighbour list overflow: %d neighbours > capacity %d", who,
                     h[F_MAXNEIGH], c->cap);
    if (e == MFA_ERR_CELL_OVERFLOW)
      return set_err(c, e, "%sbinning cell overflow: more than %d particles in a cell", who,
                     c->cell_cap);
    if (e == MFA_ERR_CELL) return set_err(c, e, "%sparticle outside the cell grid", who);
    if (e == MFA_ERR_NAN) return set_err(c, e, "%snon-finite force", who);
    return set_err(c, e, "%sdevice error %d", who, e);
  }
  return MFA_OK;
}

// A rebuild ordered by the host (init_config's binning, mfa_binning, a new
// state): clear the counts, raise the flag, rebuild, lower the flag.
static int ensure_list(mfa_ctx* c) {
  if (!c->need_rebuild) return MFA_OK;
  hipLaunchKernelGGL(k_zero_counts, dim3(grid_of(c->n_cells)), dim3(MFA_BLOCK), 0, c->stream,
                     c->n_cells, c->cell_count);
  hipLaunchKernelGGL(k_set_flag, dim3(grid_of(c->n_rep)), dim3(MFA_BLOCK), 0, c->stream, c->flags,
                     c->n_rep, (int)F_REBUILD, 1);
  int rc = launch_rebuild(c);
  if (rc) return rc;
  hipLaunchKernelGGL(k_set_flag, dim3(grid_of(c->n_rep)), dim3(MFA_BLOCK), 0, c->stream, c->flags,
                     c->n_rep, (int)F_REBUILD, 0);
  c->need_rebuild = false;
  return check_flags(c);
}

static void drop_graph(mfa_ctx* c) {
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  c->gexec = nullptr;
}

static void free_half(mfa_ctx* c) {
  if (c->pair_s) hipFree(c->pair_s);
  if (c->mir) hipFree(c->mir);
  if (c->fpart) hipFree(c->fpart);
  c->pair_s = nullptr;
  c->mir = nullptr;
  c->fpart = nullptr;
  c->half_cap = 0;
}

// Buffers of the half-list form for the current list capacity: allocated
// when the form is switched on and again when the capacity grows; kept when
// it is switched off (the autotuner switches it back and forth).
static int ensure_half(mfa_ctx* c) {
  if (!c->half_list || (c->pair_s && c->half_cap == c->cap)) return MFA_OK;
  if ((size_t)c->n * (size_t)c->cap > 0x7fffffffu)
    return set_err(c, MFA_ERR_ARG, "half-list form: %d particles x list capacity %d exceed 2^31 slots",
                   c->n, c->cap);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  free_half(c);
  const size_t slots = (size_t)c->n * (size_t)c->cap;
  if (hipMalloc((void**)&c->pair_s, slots * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&c->mir, slots * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc((void**)&c->fpart, 3 * (size_t)c->n * sizeof(double)) != hipSuccess) {
    free_half(c);
    return set_err(c, MFA_ERR_NOMEM, "hipMalloc of the half-list buffers (%zu slots) failed", slots);
  }
  c->half_cap = c->cap;
  // no mirror positions in the new buffers yet
  hipLaunchKernelGGL(k_set_flag, dim3(grid_of(c->n_rep)), dim3(MFA_BLOCK), 0, c->stream, c->flags,
                     c->n_rep, (int)F_MIRROR, 1);
  HIPCHK(c, hipGetLastError());
  return MFA_OK;
}

// One step after its verlet_positions: rebuild when due, force pass, and
// verlet_velocities fused with the next step's verlet_positions.
// new_dpd_fd (DPD_VV): the recomputed dissipative force and its application.
template <int TPP>
static void launch_dpd_fd_t(mfa_ctx* c) {
  hipLaunchKernelGGL((k_dpd_fd<TPP>), dim3(grid_of(c->n * TPP)), dim3(MFA_BLOCK), 0, c->stream,
                     c->d_kp, c->d_rmap, c->r, c->v, c->type, c->perm, c->nl, c->nn, c->cap,
                     c->fd_new);
}
static int launch_dpd_vv(mfa_ctx* c, int energy, int inc_step) {
  switch (c->tpp) {
    case 1: launch_dpd_fd_t<1>(c); break;
    case 2: launch_dpd_fd_t<2>(c); break;
    case 4: launch_dpd_fd_t<4>(c); break;
    case 8: launch_dpd_fd_t<8>(c); break;
    case 16: launch_dpd_fd_t<16>(c); break;
    default: return set_err(c, MFA_ERR_ARG, "threads per particle %d unsupported", c->tpp);
  }
  hipLaunchKernelGGL(k_dpd_vv_apply, dim3(grid_of(c->n)), dim3(MFA_BLOCK), 0, c->stream, c->d_kp,
                     c->d_rmap, c->f, c->fd, c->fd_new, c->v, c->a, c->type, c->sidx, c->vel4,
                     c->part, energy, c->d_step, inc_step);
  HIPCHK(c, hipGetLastError());
  return MFA_OK;
}

// One step after its verlet_positions: rebuild when due, force pass, and
// verlet_velocities fused with the next step's verlet_positions (with
// DPD_VV: verlet_velocities, new_dpd_fd, then the next verlet_positions).
static bool ewald_active(const mfa_ctx* c);

// Swaps the position and view buffers after a step whose force pass wrote
// the other ones.
static void swap_buffers(mfa_ctx* c) {
  std::swap(c->r, c->r_alt);
  std::swap(c->v, c->v_alt);
  std::swap(c->pos4, c->pos4_alt);
  std::swap(c->vel4, c->vel4_alt);
  c->parity ^= 1;
}

// Whether the middle steps of mfa_run_steps integrate inside the force pass:
// either form without DPD_VV (whose dissipative force needs the velocities
// of the whole half step), and not while kernels are timed one by one.
static bool fused_ok(const mfa_ctx* c) { return !c->dpd_vv && !c->timing; }

static int ew_tables(mfa_ctx* c);
static int ew_rho(mfa_ctx* c);
static int ew_sum(mfa_ctx* c);
static int ew_force(mfa_ctx* c);

static int enqueue_middle_step(mfa_ctx* c) {
  if (fused_ok(c)) {
    // rebuild when due (the step counter moves on there), then (charged
    // systems) the k-space force partials of this step's positions, then
    // the force pass with the k-space force, verlet_velocities and the next
    // verlet_positions in it
    int rc = launch_rebuild(c, nullptr, 1);
    if (rc) return rc;
    if (ewald_active(c)) {
      rc = ew_tables(c);
      if (!rc) rc = ew_rho(c);
      if (!rc) rc = ew_sum(c);
      if (!rc) rc = ew_force(c);
      if (rc) return rc;
    }
    rc = launch_force(c, TM_ALL | TM_FUSE);
    if (rc) return rc;
    swap_buffers(c);
    return MFA_OK;
  }
  int rc = launch_rebuild(c);
  if (rc) return rc;
  rc = launch_force(c, TM_ALL);
  if (rc) return rc;
  rc = launch_ewald(c, c->p.symmetry == 0);  // md_main.f90:128-132
  if (rc) return rc;
  if (!c->dpd_vv) return launch_fused(c, 0);
  const int g = grid_of(c->n);
  hipLaunchKernelGGL(k_verlet_velocities, dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->d_rmap,
                     c->f, c->v, c->a, c->type, c->part, 0, c->sidx, c->vel4, nullptr);
  rc = launch_dpd_vv(c, 0, 1);
  if (rc) return rc;
  hipLaunchKernelGGL(k_verlet_positions, dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->d_rmap,
                     c->r, c->v, c->a, c->mic, c->r_ref, c->flags, 1, c->n_cells, c->cell_count,
                     c->type, c->cell_of, c->sidx, c->pos4, c->vel4);
  HIPCHK(c, hipGetLastError());
  return MFA_OK;
}

// Captures graph_steps middle steps once per (precision, threads per particle).
static int ensure_graph(mfa_ctx* c) {
  const bool fused = fused_ok(c);
  if (c->gexec && c->gexec_precision == c->precision && c->gexec_tpp == c->tpp &&
      c->gexec_fused == fused && (!fused || c->gexec_parity == c->parity))
    return MFA_OK;
  drop_graph(c);
  hipGraph_t g = nullptr;
  const int parity0 = c->parity;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  int rc = MFA_OK;
  for (int s = 0; s < c->graph_steps && rc == MFA_OK; ++s) rc = enqueue_middle_step(c);
  c->capturing = false;
  // capturing ran nothing: the buffers are still those of parity0
  if (c->parity != parity0) swap_buffers(c);
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return set_err(c, MFA_ERR_HIP, "stream capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->gexec = nullptr;
    return set_err(c, MFA_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
  }
  c->gexec_precision = c->precision;
  c->gexec_tpp = c->tpp;
  c->gexec_fused = fused;
  c->gexec_parity = parity0;
  return MFA_OK;
}

// ewald_k(0) (ewald_k.f90:35-146): the half-space k grid inside |k| <= k_c,
// without k = 0, the lower triangle of the plane kz = 0 and half of its
// diagonal, in the reference's order (kz, ky outer, kx inner), grouped into
// runs of consecutive kx (the rows of the force kernel) and segments of at
// most EW_SEG vectors (the structure factor); slab period 3 Lz with walls.
// Host only: the plan is built, checked and then uploaded.
static void ewald_grid(const mfa_params& p, EwPlan* w) {
  const double pi_f = (double)3.1415926536f;  // mfa_common.f90:332
  double dk[3], z_slab;
  if (p.symmetry == 1) {
    for (int d = 0; d < 3; ++d) dk[d] = 2. * pi_f * p.inv_boundary[d];
    z_slab = 0.;
  } else {
    z_slab = 2.;
    dk[0] = 2. * pi_f * p.inv_boundary[0];
    dk[1] = 2. * pi_f * p.inv_boundary[1];
    dk[2] = 2. * pi_f / (p.boundary[2] + z_slab * p.boundary[2]);
  }
  const double k_c = (dk[0] + dk[1]) * (double)p.n_kmax / 2.;
  const int nxm = (int)(k_c / dk[0]), nym = (int)(k_c / dk[1]), nzm = (int)(k_c / dk[2]);
  const double fac_alpha = 1. / (4. * p.alpha_c);
  w->kv.clear();
  w->rows.clear();
  w->row_k0.clear();
  for (int nkz = 0; nkz <= nzm; ++nkz)
    for (int nky = -nym; nky <= nym; ++nky) {
      int open = -1;  // index of the open row
      for (int nkx = -nxm; nkx <= nxm; ++nkx) {
        const double ax = (double)nkx * dk[0], ay = (double)nky * dk[1], az = (double)nkz * dk[2];
        const double rk_2 = ax * ax + ay * ay + az * az;
        bool take = !(rk_2 > k_c * k_c);
        if (nkz == 0 && nky == 0 && nkx == 0) take = false;
        if (nkz == 0 && nkx + nky < 0) take = false;
        if (nkz == 0 && -nkx == nky && nky > 0) take = false;
        if (!take) {
          open = -1;
          continue;
        }
        const double kx = dk[0] * (double)nkx, ky = dk[1] * (double)nky, kz = dk[2] * (double)nkz;
        const double k2 = kx * kx + ky * ky + kz * kz;
        w->kv.push_back(make_double4(kx, ky, kz, (1. / k2) * exp(-k2 * fac_alpha)));
        if (open < 0) {
          w->rows.push_back(make_int4(nky, nkz, nkx, nkx));
          w->row_k0.push_back((int)w->kv.size() - 1);
          open = (int)w->rows.size() - 1;
        } else {
          w->rows[(size_t)open].w = nkx;
        }
      }
    }
  // segments of at most EW_SEG vectors, longest first (stable), so the
  // lanes of a wave loop over equal lengths
  std::vector<int4> segs;
  std::vector<int> seg_k0;
  for (size_t q = 0; q < w->rows.size(); ++q)
    for (int lo = w->rows[q].z; lo <= w->rows[q].w; lo += EW_SEG) {
      const int len = w->rows[q].w - lo + 1 < EW_SEG ? w->rows[q].w - lo + 1 : EW_SEG;
      segs.push_back(make_int4(w->rows[q].x, w->rows[q].y, lo, len));
      seg_k0.push_back(w->row_k0[q] + (lo - w->rows[q].z));
    }
  std::vector<int> order(segs.size());
  for (size_t q = 0; q < order.size(); ++q) order[q] = (int)q;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return segs[a].w > segs[b].w; });
  w->segs.resize(segs.size());
  w->seg_k0.resize(segs.size());
  for (size_t q = 0; q < order.size(); ++q) {
    w->segs[q] = segs[(size_t)order[q]];
    w->seg_k0[q] = seg_k0[(size_t)order[q]];
  }
  EwaldK& e = w->e;
  memset(&e, 0, sizeof e);
  for (int d = 0; d < 3; ++d) e.dk[d] = dk[d];
  e.lz_half = p.boundary[2] / 2.;
  e.nk = (int)w->kv.size();
  e.nsegs = (int)w->segs.size();
  e.nxm = nxm;
  e.nym = nym;
  e.nzm = nzm;
  e.nex = 2 * nxm + 1;
  e.ney = 2 * nym + 1;
  e.nez = nzm + 1;
  w->inv_V = p.inv_boundary[0] * p.inv_boundary[1] / (p.boundary[2] * (1. + z_slab));
  w->sum_blocks = (e.nk + 63) / 64;
  w->rows_kernel = false;
  w->rows_per_group = w->n_kgroups = w->force_blocks = w->force_grid = 0;
}

// Launch shapes for nq charged particles out of n_alloc (the padded particle
// count the buffers are sized for), and every bound the kernels rely on:
// LDS within lds_limit, table and row indices inside their tables, the force
// partials inside their buffer, the XCD-aware grid a multiple of 8 whose
// decoding covers exactly the (particle block, row group) pairs.  rows_ok:
// the row kernel exists for this grid and runs without register spills.
// Returns 0, or MFA_ERR_ARG with the reason in *why.
static int ewald_shape(EwPlan* w, int nq, size_t n_alloc, size_t lds_limit, bool rows_ok,
                       std::string* why) {
  EwaldK& e = w->e;
  char buf[256];
  auto fail = [&](const char* fmt, long long a, long long b) {
    snprintf(buf, sizeof buf, fmt, a, b);
    *why = buf;
    return MFA_ERR_ARG;
  };
  e.nq = nq;
  w->fk_cap = EW_FK_CAP(n_alloc);
  w->lds_rho = (size_t)EW_CH * (e.nex + e.ney + e.nez + EW_SEG) * sizeof(double2);
  if (w->lds_rho > lds_limit)
    return fail("Ewald: structure-factor tile of %lld bytes exceeds the LDS limit %lld",
                (long long)w->lds_rho, (long long)lds_limit);
  if ((size_t)nq > n_alloc) return fail("Ewald: %lld charges > %lld particles", nq, (long long)n_alloc);
  // every row inside the grid and the k index range; rows are the maximal
  // runs, so they cover the k vectors in order
  long long next_k = 0;
  for (size_t q = 0; q < w->rows.size(); ++q) {
    const int4 r = w->rows[q];
    if (r.z < -e.nxm || r.w > e.nxm || r.w < r.z || r.x < -e.nym || r.x > e.nym || r.y < 0 ||
        r.y > e.nzm)
      return fail("Ewald: row %lld outside the k grid (nx_hi %lld)", (long long)q, r.w);
    if (w->row_k0[q] != next_k) return fail("Ewald: row %lld starts at k %lld", (long long)q, w->row_k0[q]);
    next_k += r.w - r.z + 1;
  }
  if (next_k != e.nk) return fail("Ewald: rows hold %lld of %lld k vectors", next_k, e.nk);
  for (size_t q = 0; q < w->segs.size(); ++q)
    if (w->seg_k0[q] < 0 || w->seg_k0[q] + w->segs[q].w > e.nk || w->segs[q].w < 1 ||
        w->segs[q].w > EW_SEG)
      return fail("Ewald: segment %lld outside the k range (k0 %lld)", (long long)q, w->seg_k0[q]);
  w->rows_kernel = rows_ok && !w->rows.empty() && e.nxm >= 1 && e.nxm <= EW_ROWS_MAX_NXM;
  if (nq == 0 || e.nsegs == 0) return MFA_OK;
  e.n_chunks = (nq + EW_CH - 1) / EW_CH;
  e.n_mz = (nq + MFA_BLOCK - 1) / MFA_BLOCK;
  if (w->rows_kernel) {
    // row groups: about 1024 blocks of 256 EW_PPL particles (four resident
    // waves per SIMD), each group's G rows within the LDS limit, the
    // partials within their buffer
    const int nrows = (int)w->rows.size();
    const int pb = (nq + MFA_BLOCK * EW_PPL - 1) / (MFA_BLOCK * EW_PPL);
    int ng = (1024 + pb - 1) / pb;
    const int cap = (int)(w->fk_cap / (size_t)nq);
    ng = ng > cap ? cap : ng;
    ng = ng > nrows ? nrows : (ng < 1 ? 1 : ng);
    int rpg = (nrows + ng - 1) / ng;
    const int rpg_lds = (int)(lds_limit / ((size_t)e.nex * sizeof(double2)));
    if (rpg > rpg_lds) rpg = rpg_lds;
    if (rpg < 1) return fail("Ewald: one k row of %lld entries exceeds the LDS limit %lld", e.nex,
                             (long long)lds_limit);
    w->rows_per_group = rpg;
    w->n_kgroups = (nrows + rpg - 1) / rpg;
    w->lds_force = (size_t)rpg * e.nex * sizeof(double2);
    w->force_blocks = pb * w->n_kgroups;
    w->force_grid = (w->force_blocks + 7) / 8 * 8;
    // the kernel decodes block b as L = (b % 8) (grid / 8) + b / 8, particle
    // block L / groups, group L % groups, and skips L >= force_blocks
    if (w->force_grid % 8 != 0 || (long long)(pb - 1) * MFA_BLOCK * EW_PPL >= nq ||
        (w->n_kgroups - 1) * rpg >= nrows)
      return fail("Ewald: inconsistent row launch (%lld blocks, %lld groups)", w->force_blocks,
                  w->n_kgroups);
  } else {
    // segment kernel: about 8192 waves (4 per block), block rows of segments
    const int qb = (nq + 63) / 64;
    int kg = 2048 / qb;
    kg = kg < 1 ? 1 : (kg > e.nsegs ? e.nsegs : kg);
    e.segs_per_kgroup = (e.nsegs + kg - 1) / kg;
    w->n_kgroups = (e.nsegs + e.segs_per_kgroup - 1) / e.segs_per_kgroup;
    w->lds_force = ((size_t)e.nex * 64 + (size_t)e.segs_per_kgroup * EW_SEG) * sizeof(double2);
    if (w->lds_force > lds_limit)
      return fail("Ewald: force tile of %lld bytes exceeds the LDS limit %lld", (long long)w->lds_force,
                  (long long)lds_limit);
  }
  if ((size_t)w->n_kgroups * (size_t)nq > w->fk_cap)
    return fail("Ewald: %lld force partials exceed the buffer of %lld",
                (long long)w->n_kgroups * nq, (long long)w->fk_cap);
  return MFA_OK;
}

// Whether k_ewald_force_rows<nxm> runs without scratch (register spills).
static bool rows_kernel_fits(int nxm) {
  hipFuncAttributes a;
  hipError_t err = hipErrorInvalidValue;
#define EW_ROWS_ATTR(X) \
  case X: err = hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_ewald_force_rows<X>)); break;
  switch (nxm) {
    EW_ROWS_ATTR(1) EW_ROWS_ATTR(2) EW_ROWS_ATTR(3) EW_ROWS_ATTR(4) EW_ROWS_ATTR(5) EW_ROWS_ATTR(6)
    EW_ROWS_ATTR(7) EW_ROWS_ATTR(8) EW_ROWS_ATTR(9) EW_ROWS_ATTR(10) EW_ROWS_ATTR(11) EW_ROWS_ATTR(12)
    default: return false;
  }
#undef EW_ROWS_ATTR
  return err == hipSuccess && a.localSizeBytes == 0;
}

// Device side of ewald_k(0): the plan of the box, its buffers and tables.
static int ewald_setup(mfa_ctx* c) {
  EwPlan& w = c->ew;
  ewald_grid(c->p, &w);
  const EwaldK& e = w.e;
  c->ew_rows_ok = e.nxm >= 1 && e.nxm <= EW_ROWS_MAX_NXM && rows_kernel_fits(e.nxm);
  std::string why;
  if (ewald_shape(&w, 0, (size_t)c->n, c->lds_limit, c->ew_rows_ok, &why))
    return set_err(nullptr, MFA_ERR_ARG, "%s", why.c_str());
  const size_t n = (size_t)c->n;
  const size_t n_chunk_max = (n + EW_CH - 1) / EW_CH;
  const size_t nk = (size_t)(e.nk > 0 ? e.nk : 1), nr = (size_t)(e.nsegs > 0 ? e.nsegs : 1);
  const size_t nrw = w.rows.empty() ? 1 : w.rows.size();
  const size_t ntab = (size_t)(e.nex + e.ney + e.nez);
  void** bufs[] = {(void**)&c->d_kvec, (void**)&c->d_runs, (void**)&c->d_run_k0,
                   (void**)&c->d_qidx, (void**)&c->d_qval, (void**)&c->d_rho_part,
                   (void**)&c->d_grho, (void**)&c->d_uk_part, (void**)&c->d_mz_part,
                   (void**)&c->d_mz, (void**)&c->d_fk_part, (void**)&c->d_ek, (void**)&c->d_tx,
                   (void**)&c->d_rows, (void**)&c->d_row_k0, (void**)&c->d_qinv};
  const size_t bytes[] = {nk * sizeof(double4), nr * sizeof(int4), nr * sizeof(int),
                          n * sizeof(int), n * sizeof(double), n_chunk_max * nk * sizeof(double2),
                          nk * sizeof(double2), (size_t)(w.sum_blocks + 1) * sizeof(double),
                          (n / MFA_BLOCK + 1) * sizeof(double), sizeof(double),
                          w.fk_cap * 3 * sizeof(double), 2 * sizeof(double),
                          ntab * n * sizeof(double2), nrw * sizeof(int4), nrw * sizeof(int), n * sizeof(int)};
  for (int b = 0; b < 16; ++b)
    if (hipMalloc(bufs[b], bytes[b]) != hipSuccess)
      return set_err(nullptr, MFA_ERR_NOMEM, "hipMalloc of %zu bytes (Ewald) failed", bytes[b]);
  c->d_ty = c->d_tx + (size_t)e.nex * n;
  c->d_tz = c->d_ty + (size_t)e.ney * n;
  if (e.nk > 0) {
    HIPCHK(nullptr, hipMemcpy(c->d_kvec, w.kv.data(), w.kv.size() * sizeof(double4), hipMemcpyHostToDevice));
    HIPCHK(nullptr, hipMemcpy(c->d_runs, w.segs.data(), w.segs.size() * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHK(nullptr, hipMemcpy(c->d_run_k0, w.seg_k0.data(), w.seg_k0.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(nullptr, hipMemcpy(c->d_rows, w.rows.data(), w.rows.size() * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHK(nullptr, hipMemcpy(c->d_row_k0, w.row_k0.data(), w.row_k0.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  HIPCHK(nullptr, hipMemset(c->d_mz, 0, sizeof(double)));
  c->ewald = true;
  return MFA_OK;
}

// Charged particles of a new state (q(i) = charge(a_type(i)),
// init_config.f90:221-226), the self term (:235-239) and launch shapes.
static int ewald_set_charges(mfa_ctx* c, const int32_t* a_type) {
  const double pi_f = (double)3.1415926536f;
  std::vector<int> qi;
  std::vector<double> qv;
  double self = 0.;
  for (int i = 0; i < c->p.n_part; ++i) {
    const double q = c->p.charge[a_type[i] - 1];
    self = self - q * q;
    if (q != 0.0) {
      qi.push_back(i);
      qv.push_back(q);
    }
  }
  c->v_coul_self = self * c->p.sqrt_alpha / sqrt(pi_f);
  std::string why;
  if (ewald_shape(&c->ew, (int)qi.size(), (size_t)c->n, c->lds_limit, c->ew_rows_ok, &why)) {
    c->ew.e.nq = 0;  // no launch of an unchecked shape
    return set_err(c, MFA_ERR_ARG, "%s", why.c_str());
  }
  if (qi.empty() || c->ew.e.nsegs == 0) return MFA_OK;
  HIPCHK(c, hipMemcpy(c->d_qidx, qi.data(), qi.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_qval, qv.data(), qv.size() * sizeof(double), hipMemcpyHostToDevice));
  std::vector<int> qinv((size_t)c->n, -1);
  for (size_t q = 0; q < qi.size(); ++q) qinv[(size_t)qi[q]] = (int)q;
  HIPCHK(c, hipMemcpy(c->d_qinv, qinv.data(), qinv.size() * sizeof(int), hipMemcpyHostToDevice));
  return MFA_OK;
}

static bool ewald_active(const mfa_ctx* c) { return c->ewald && c->ew.e.nq > 0 && c->ew.e.nsegs > 0; }

// k-space forces (and, with add_dip, the slab dipole forces) added to F.
// Shapes come from the checked plan (ewald_shape).
// The five kernels of the k-space stage, one function each (the timing
// helper launches them one at a time).  Shapes come from the checked plan
// (ewald_shape).
static int ew_tables(mfa_ctx* c) {
  const EwaldK& e = c->ew.e;
  hipLaunchKernelGGL(k_ewald_tables, dim3(e.n_mz, 3), dim3(MFA_BLOCK), 0, c->stream, e, c->r,
                     c->d_qidx, c->d_qval, c->d_tx, c->d_ty, c->d_tz, c->d_mz_part);
  return MFA_OK;
}
static int ew_rho(mfa_ctx* c) {
  const EwPlan& w = c->ew;
  const EwaldK& e = w.e;
  hipLaunchKernelGGL(k_ewald_rho, dim3(e.n_chunks, (e.nsegs + MFA_BLOCK - 1) / MFA_BLOCK),
                     dim3(MFA_BLOCK), w.lds_rho, c->stream, e, c->d_qval, c->d_tx, c->d_ty, c->d_tz,
                     c->d_runs, c->d_run_k0, c->d_rho_part);
  return MFA_OK;
}
static int ew_sum(mfa_ctx* c) {
  const EwPlan& w = c->ew;
  hipLaunchKernelGGL(k_ewald_sum, dim3(w.sum_blocks), dim3(64 * EW_SUMW), 0, c->stream, w.e,
                     c->d_rho_part, c->d_kvec, c->d_grho, c->d_uk_part, c->d_mz_part, c->d_mz);
  return MFA_OK;
}
static int ew_force(mfa_ctx* c) {
  const EwPlan& w = c->ew;
  const EwaldK& e = w.e;
  if (w.rows_kernel) {
#define EW_ROWS_CASE(X)                                                                              \
  case X:                                                                                            \
    hipLaunchKernelGGL((k_ewald_force_rows<X>), dim3(w.force_grid), dim3(MFA_BLOCK), w.lds_force,    \
                       c->stream, e, c->d_tx, c->d_ty, c->d_tz, c->d_rows, c->d_row_k0,             \
                       (int)w.rows.size(), w.rows_per_group, c->d_grho, c->d_fk_part);              \
    break;
    switch (e.nxm) {
      EW_ROWS_CASE(1) EW_ROWS_CASE(2) EW_ROWS_CASE(3) EW_ROWS_CASE(4) EW_ROWS_CASE(5) EW_ROWS_CASE(6)
      EW_ROWS_CASE(7) EW_ROWS_CASE(8) EW_ROWS_CASE(9) EW_ROWS_CASE(10) EW_ROWS_CASE(11) EW_ROWS_CASE(12)
      default: return set_err(c, MFA_ERR_ARG, "Ewald: no row kernel for nxm = %d", e.nxm);
    }
#undef EW_ROWS_CASE
  } else {
    hipLaunchKernelGGL(k_ewald_force, dim3((e.nq + 63) / 64, w.n_kgroups), dim3(MFA_BLOCK),
                       w.lds_force, c->stream, e, c->d_tx, c->d_ty, c->d_tz, c->d_runs, c->d_run_k0,
                       c->d_grho, c->d_fk_part);
  }
  return MFA_OK;
}
static int ew_apply(mfa_ctx* c, int add_dip) {
  const EwPlan& w = c->ew;
  const EwaldK& e = w.e;
  hipLaunchKernelGGL(k_ewald_apply, dim3((3 * e.nq + MFA_BLOCK - 1) / MFA_BLOCK), dim3(MFA_BLOCK), 0,
                     c->stream, e, c->d_qidx, c->d_qval, c->d_fk_part, w.n_kgroups, c->d_mz,
                     w.inv_V, 1, add_dip, c->f);
  return MFA_OK;
}
static int ew_apply_stage(mfa_ctx* c) { return ew_apply(c, c->p.symmetry == 0 ? 1 : 0); }

// k-space forces (and, with add_dip, the slab dipole forces) added to F.
static int launch_ewald(mfa_ctx* c, int add_dip) {
  if (!ewald_active(c)) return MFA_OK;
  int rc = ew_tables(c);
  if (!rc) rc = ew_rho(c);
  if (!rc) rc = ew_sum(c);
  if (!rc) rc = ew_force(c);
  if (!rc) rc = ew_apply(c, add_dip);
  if (rc) return rc;
  HIPCHK(c, hipGetLastError());
  return MFA_OK;
}

// Host-only: the k-space plan of a box for n_charged charges (see the
// header for the fields).  The row kernel's register check needs the device
// and runs at mfa_create; here the row kernel is assumed to fit.
extern "C" int32_t mfa_ewald_plan(const mfa_params* p, int32_t n_charged, int64_t* out,
                                  int32_t n_out) {
  if (!p || !out || n_out < MFA_EWPLAN_N || n_charged < 0 || n_charged > p->n_part)
    return set_err(nullptr, MFA_ERR_ARG, "mfa_ewald_plan: bad argument");
  EwPlan w;
  ewald_grid(*p, &w);
  const size_t n_alloc = ((size_t)p->n_part + MFA_BLOCK - 1) / MFA_BLOCK * MFA_BLOCK;
  std::string why;
  const bool rows_ok = w.e.nxm >= 1 && w.e.nxm <= EW_ROWS_MAX_NXM;
  const int rc = ewald_shape(&w, n_charged, n_alloc, 65536, rows_ok, &why);
  const int64_t v[MFA_EWPLAN_N] = {w.e.nk, w.e.nsegs, (int64_t)w.rows.size(), w.e.nxm, w.e.nym,
                                   w.e.nzm, w.rows_kernel ? 1 : 0, w.rows_per_group, w.n_kgroups,
                                   w.force_blocks, w.force_grid, (int64_t)w.lds_rho,
                                   (int64_t)w.lds_force, (int64_t)w.n_kgroups * n_charged,
                                   (int64_t)w.fk_cap, w.e.n_chunks};
  for (int q = 0; q < MFA_EWPLAN_N; ++q) out[q] = v[q];
  if (rc) return set_err(nullptr, rc, "%s", why.c_str());
  return MFA_OK;
}

// u_cou_k and the dipole energy of the last launch_ewald / dipole update.
static int ewald_energies(mfa_ctx* c, double* uk, double* udip) {
  *uk = *udip = 0.0;
  if (!ewald_active(c)) return MFA_OK;
  hipLaunchKernelGGL(k_ewald_energy, dim3(1), dim3(64), 0, c->stream, c->d_uk_part, c->ew.sum_blocks,
                     c->d_mz, c->ew.inv_V, c->d_ek);
  HIPCHK(c, hipGetLastError());
  double h[2];
  HIPCHK(c, hipMemcpyAsync(h, c->d_ek, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *uk = h[0];
  *udip = c->p.symmetry == 0 ? h[1] : 0.0;
  return MFA_OK;
}

// ---------------------------------------------------------------------------
// C ABI

extern "C" const char* mfa_last_error(mfa_ctx* c) {
  return c ? c->err.c_str() : g_noctx_err.c_str();
}

// One context for n_rep replicas (n_rep = 1 for a single system).  Replica q
// owns particles [p0, p0 + n_part) with p0 a multiple of 256 and cells
// [c0, c0 + n_cells); cell and list capacities are the largest any replica
// needs.  Replicas share the compile-time switches of the reference (system,
// symmetry, thermostat, DPD cut-off; bending may differ) and carry no charges.
static int32_t create_impl(const mfa_params* ps, int32_t n_rep, int32_t precision,
                           int32_t list_capacity, mfa_ctx** out) {
  if (!ps || !out || n_rep < 1) return set_err(nullptr, MFA_ERR_ARG, "null argument");
  if (precision != MFA_FP64 && precision != MFA_MIXED)
    return set_err(nullptr, MFA_ERR_ARG, "unknown precision %d", precision);
  for (int q = 0; q < n_rep; ++q) {
    const mfa_params& p = ps[q];
    if (p.n_part <= 0) return set_err(nullptr, MFA_ERR_ARG, "replica %d: n_part must be positive", q);
    if (n_rep > 1 && (p.system != ps[0].system || p.symmetry != ps[0].symmetry ||
                      p.thermostat != ps[0].thermostat || p.dpd_cut_off_on != ps[0].dpd_cut_off_on ||
                      p.dpd_vv != ps[0].dpd_vv))
      return set_err(nullptr, MFA_ERR_ARG, "replica %d: switches differ from replica 0", q);
    if (n_rep > 1 && p.coulomb)
      return set_err(nullptr, MFA_ERR_ARG, "charged systems run one per context");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_err(nullptr, MFA_ERR_HIP, "no HIP device");
  mfa_ctx* c = new mfa_ctx();
  {
    // LDS per workgroup: the device's limit, at most the 64 KB the kernels'
    // tiles are sized for
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        lds <= 0) {
      delete c;
      return set_err(nullptr, MFA_ERR_HIP, "cannot read the device's LDS limit");
    }
    c->lds_limit = (size_t)(lds < 65536 ? lds : 65536);
  }
  c->p = ps[0];
  c->n_rep = n_rep;
  c->extra_terms = false;
  c->dpd_vv = ps[0].dpd_vv != 0;
  c->lgv_terms = false;
  for (int q = 0; q < n_rep; ++q) {
    c->extra_terms |= ps[q].coulomb != 0 || ps[q].bending != 0;
    c->lgv_terms |= ps[q].thermostat == 1;
  }
  c->ps.assign(ps, ps + n_rep);
  c->precision = precision;
  c->tpp = 8;
  c->pair_queue = 0;
  c->half_list = false;
  c->list_kind = LK_PLAIN;
  c->need_rebuild = true;
  c->timing = false;
  c->capturing = false;
  c->graph_steps = 16;
  if (const char* e = getenv("MFA_GRAPH_STEPS")) c->graph_steps = atoi(e);  // 0: direct launches
  if (c->graph_steps < 0 || c->graph_steps > 1024) c->graph_steps = 16;
  c->gexec = nullptr;
  // capacities and layout
  int cap = list_capacity, cell_cap = 0;
  int64_t ntot = 0, ctot = 0, gtot = 0;
  for (int q = 0; q < n_rep; ++q) {
    const mfa_params& p = ps[q];
    const double vol = p.boundary[0] * p.boundary[1] * p.boundary[2];
    const double dens = (double)p.n_part / vol;
    if (list_capacity <= 0) {
      const int cq = list_cap_of(p);
      cap = cq > cap ? cq : cap;
    }
    // cell capacity: twice the mean occupancy plus 32 (the reference allows 300)
    const double occ = dens * p.r_bin[0] * p.r_bin[1] * p.r_bin[2];
    const int cc = (((int)(2.0 * occ + 32.0)) + 31) / 32 * 32;
    cell_cap = cc > cell_cap ? cc : cell_cap;
    KParams k = make_kparams(p);
    k.p0 = (int)ntot;
    k.c0 = (int)ctot;
    k.fg0 = (int)gtot;
    k.rep = q;
    c->ks.push_back(k);
    const int64_t nbq = ((int64_t)p.n_part + MFA_BLOCK - 1) / MFA_BLOCK;
    ntot += nbq * MFA_BLOCK;
    gtot += (nbq + FILL_GROUP - 1) / FILL_GROUP;
    ctot += (int64_t)p.n_bin[0] * p.n_bin[1] * p.n_bin[2];
  }
  if (ntot > (int64_t)MFA_JMASK) {
    delete c;
    return set_err(nullptr, MFA_ERR_ARG, "at most %u particles per context", MFA_JMASK);
  }
  for (KParams& k : c->ks) k.cell_cap = cell_cap;
  {
    // octant groups pay off where cells hold many particles
    int64_t np = 0, nc = 0;
    for (int q = 0; q < n_rep; ++q) {
      np += ps[q].n_part;
      nc += (int64_t)ps[q].n_bin[0] * ps[q].n_bin[1] * ps[q].n_bin[2];
    }
    c->lb_octants = nc > 0 && np >= (int64_t)LB_MIN_OCTANT * nc;
  }
  if (cap > MFA_CAP_MAX) {  // row counts are packed in 16 bits
    delete c;
    return set_err(nullptr, MFA_ERR_ARG, "list capacity %d beyond %d", cap, MFA_CAP_MAX);
  }
  c->cap = cap;
  c->cap_fixed = list_capacity > 0;
  c->cell_cap = cell_cap;
  c->n = (int)ntot;
  c->n_cells = (int)ctot;
  {
    // the list build gives each particle of a cell a lane of one of its waves
    if (c->cell_cap > LB_MAX_CELL) {
      delete c;
      return set_err(nullptr, MFA_ERR_ARG, "binning cells too crowded for the list build (capacity %d)",
                     cell_cap);
    }
    if (list_lds_bytes(c) > c->lds_limit) {
      const size_t b = list_lds_bytes(c);
      delete c;
      return set_err(nullptr, MFA_ERR_ARG, "list build needs %zu bytes of LDS", b);
    }
  }
  const size_t n = (size_t)c->n;
  const size_t ncs = (size_t)c->n_cells * c->cell_cap;
  c->max_blocks = grid_of(c->n * 16) + 1;
#define ALLOC(ptr, bytes)                                                        \
  do {                                                                           \
    if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess) {                      \
      mfa_destroy(c);                                                            \
      return set_err(nullptr, MFA_ERR_NOMEM, "hipMalloc of %zu bytes failed",    \
                     (size_t)(bytes));                                           \
    }                                                                            \
  } while (0)
  ALLOC(c->r, n * 3 * sizeof(double));
  ALLOC(c->v, n * 3 * sizeof(double));
  ALLOC(c->a, n * 3 * sizeof(double));
  ALLOC(c->f, n * 3 * sizeof(double));
  ALLOC(c->r_ref, n * 3 * sizeof(double));
  ALLOC(c->type, n * sizeof(int32_t));
  ALLOC(c->mic, n * 3 * sizeof(int32_t));
  ALLOC(c->cell_of, n * sizeof(int32_t));
  ALLOC(c->cell_count, (size_t)c->n_cells * sizeof(int32_t));
  ALLOC(c->cell_slots, ncs * sizeof(int32_t));
  ALLOC(c->cell_start, ((size_t)c->n_cells + 1) * sizeof(int32_t));
  ALLOC(c->perm, n * sizeof(int32_t));
  ALLOC(c->fill_cnt, (size_t)gtot * sizeof(int32_t));
  if (c->dpd_vv) {
    ALLOC(c->fd, n * 3 * sizeof(double));
    ALLOC(c->fd_new, n * 3 * sizeof(double));
  }
  ALLOC(c->sidx, n * sizeof(int32_t));
  ALLOC(c->nl, n * (size_t)c->cap * sizeof(unsigned));
  ALLOC(c->nn, n * sizeof(int32_t));
  ALLOC(c->pos4, n * sizeof(float4));
  ALLOC(c->vel4, n * sizeof(float4));
  ALLOC(c->r_alt, n * 3 * sizeof(double));
  ALLOC(c->v_alt, n * 3 * sizeof(double));
  ALLOC(c->pos4_alt, n * sizeof(float4));
  ALLOC(c->vel4_alt, n * sizeof(float4));
  ALLOC(c->flags, (size_t)n_rep * F_NFLAGS * sizeof(int32_t));
  ALLOC(c->d_step, sizeof(long long));
  ALLOC(c->part, (size_t)c->max_blocks * MFA_NTERMS * sizeof(double));
  ALLOC(c->red, (size_t)n_rep * MFA_NTERMS * sizeof(double));
  ALLOC(c->d_kp, (size_t)n_rep * sizeof(KParams));
  ALLOC(c->d_rmap, (n / MFA_BLOCK + 1) * sizeof(int));
  ALLOC(c->d_cell_rep, (size_t)c->n_cells * sizeof(int));
#undef ALLOC
  if (hipHostMalloc((void**)&c->h_flags, (size_t)n_rep * F_NFLAGS * sizeof(int32_t),
                    hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_red, (size_t)n_rep * MFA_NTERMS * sizeof(double),
                    hipHostMallocDefault) != hipSuccess) {
    mfa_destroy(c);
    return set_err(nullptr, MFA_ERR_NOMEM, "hipHostMalloc failed");
  }
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    mfa_destroy(c);
    return set_err(nullptr, MFA_ERR_HIP, "hipStreamCreate failed");
  }
  {
    std::vector<int> rmap(n / MFA_BLOCK + 1, 0), crep((size_t)c->n_cells, 0);
    for (int q = 0; q < n_rep; ++q) {
      const KParams& k = c->ks[(size_t)q];
      const int nb = grid_of(k.n);
      for (int b = 0; b < nb; ++b) rmap[(size_t)(k.p0 / MFA_BLOCK + b)] = q;
      for (int cc = 0; cc < k.n_cells; ++cc) crep[(size_t)(k.c0 + cc)] = q;
    }
    if (hipMemcpy(c->d_kp, c->ks.data(), (size_t)n_rep * sizeof(KParams), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(c->d_rmap, rmap.data(), rmap.size() * sizeof(int), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(c->d_cell_rep, crep.data(), crep.size() * sizeof(int), hipMemcpyHostToDevice) !=
            hipSuccess) {
      mfa_destroy(c);
      return set_err(nullptr, MFA_ERR_HIP, "upload of the replica tables failed");
    }
  }
  c->rep_obs.assign((size_t)n_rep, mfa_obs());
  c->ev_force_used = c->ev_integ_used = 0;
  if (c->p.coulomb) {
    const int rc = ewald_setup(c);
    if (rc) {
      mfa_destroy(c);
      return rc;
    }
  }
  // padding particles stay type 1 at the origin and never interact
  std::vector<int32_t> ones(n, 1);
  hipMemcpyAsync(c->type, ones.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
  hipMemsetAsync(c->r, 0, n * 3 * sizeof(double), c->stream);
  hipMemsetAsync(c->r_alt, 0, n * 3 * sizeof(double), c->stream);
  hipMemsetAsync(c->v_alt, 0, n * 3 * sizeof(double), c->stream);
  hipMemsetAsync(c->v, 0, n * 3 * sizeof(double), c->stream);
  hipMemsetAsync(c->r_ref, 0, n * 3 * sizeof(double), c->stream);
  hipMemsetAsync(c->cell_count, 0, (size_t)c->n_cells * sizeof(int32_t), c->stream);
  hipMemsetAsync(c->flags, 0, (size_t)n_rep * F_NFLAGS * sizeof(int32_t), c->stream);
  hipMemsetAsync(c->d_step, 0, sizeof(long long), c->stream);
  hipMemsetAsync(c->fill_cnt, 0, (size_t)gtot * sizeof(int32_t), c->stream);
  hipMemsetAsync(c->mic, 0, n * 3 * sizeof(int32_t), c->stream);
  hipMemsetAsync(c->f, 0, n * 3 * sizeof(double), c->stream);
  hipMemsetAsync(c->a, 0, n * 3 * sizeof(double), c->stream);
  if (c->dpd_vv) {
    hipMemsetAsync(c->fd, 0, n * 3 * sizeof(double), c->stream);
    hipMemsetAsync(c->fd_new, 0, n * 3 * sizeof(double), c->stream);
  }
  hipMemsetAsync(c->nn, 0, n * sizeof(int32_t), c->stream);
  hipMemsetAsync(c->cell_of, 0, n * sizeof(int32_t), c->stream);
  {
    // identity order until the first rebuild (padding keeps it for good)
    std::vector<int32_t> id(n);
    for (size_t q = 0; q < n; ++q) id[q] = (int32_t)q;
    hipMemcpyAsync(c->perm, id.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
    hipMemcpyAsync(c->sidx, id.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    mfa_destroy(c);
    return set_err(nullptr, MFA_ERR_HIP, "initial clears failed");
  }
  *out = c;
  return MFA_OK;
}

extern "C" int32_t mfa_create(const mfa_params* p, int32_t precision, int32_t list_capacity,
                              mfa_ctx** out) {
  return create_impl(p, 1, precision, list_capacity, out);
}

extern "C" int32_t mfa_create_batch(const mfa_params* params, int32_t n_rep, int32_t precision,
                                    int32_t list_capacity, mfa_ctx** out) {
  return create_impl(params, n_rep, precision, list_capacity, out);
}

extern "C" int32_t mfa_destroy(mfa_ctx* c) {
  if (!c) return MFA_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  drop_graph(c);
  free_half(c);
  void* ptrs[] = {c->r,   c->v,       c->a,          c->f,          c->r_ref, c->type,
                  c->mic, c->cell_of, c->cell_count, c->cell_slots, c->nl,    c->nn,
                  c->pos4, c->vel4,   c->flags,      c->d_step,     c->part,  c->red,
                  c->d_kp, c->d_rmap, c->d_cell_rep, c->cell_start, c->perm,  c->sidx,
                  c->fd,   c->fd_new, c->fill_cnt, c->head_f, c->d_hist, c->d_chain_rg,
                  c->d_obs, c->r_alt, c->v_alt, c->pos4_alt, c->vel4_alt};
  for (void* q : ptrs)
    if (q) hipFree(q);
  if (c->d_rows) hipFree(c->d_rows);
  if (c->d_row_k0) hipFree(c->d_row_k0);
  if (c->d_qinv) hipFree(c->d_qinv);
  void* eptrs[] = {c->d_kvec,    c->d_runs,    c->d_run_k0, c->d_qidx, c->d_qval,   c->d_rho_part,
                   c->d_grho,    c->d_uk_part, c->d_mz_part, c->d_mz,  c->d_fk_part, c->d_ek,
                   c->d_tx};
  for (void* q : eptrs)
    if (q) hipFree(q);
  if (c->h_flags) hipHostFree(c->h_flags);
  if (c->h_red) hipHostFree(c->h_red);
  for (hipEvent_t e : c->ev_force) hipEventDestroy(e);
  for (hipEvent_t e : c->ev_integ) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return MFA_OK;
}

// The row layout the builds make: half-list rows exactly when the half-list
// passes will read them (half-list form on, mixed form, 4 or more threads
// per particle), near/shell rows for the pair-queue mode that queues the skin
// part, plain rows otherwise.  Every pass walks both parts of a row (the
// FP64, DPD_VV and adjoint passes through list_slot).  A change of layout
// takes a fresh build before the next pass (the pair set is the same).
static void sync_split(mfa_ctx* c) {
  const bool mixed = c->precision == MFA_MIXED;
  const int want = c->half_list && mixed && c->tpp >= 4 ? LK_HALF
                   : mixed && c->pair_queue == 2         ? LK_NEAR
                                                         : LK_PLAIN;
  if (want != c->list_kind) {
    c->list_kind = want;
    c->need_rebuild = true;
    drop_graph(c);
  }
}

extern "C" int32_t mfa_set_precision(mfa_ctx* c, int32_t precision) {
  if (!c) return MFA_ERR_ARG;
  if (precision != MFA_FP64 && precision != MFA_MIXED)
    return set_err(c, MFA_ERR_ARG, "unknown precision %d", precision);
  c->precision = precision;
  sync_split(c);
  return MFA_OK;
}

extern "C" int32_t mfa_set_threads_per_particle(mfa_ctx* c, int32_t tpp) {
  if (!c) return MFA_ERR_ARG;
  if (tpp != 1 && tpp != 2 && tpp != 4 && tpp != 8 && tpp != 16)
    return set_err(c, MFA_ERR_ARG, "threads per particle must be 1,2,4,8 or 16");
  c->tpp = tpp;
  sync_split(c);
  return MFA_OK;
}

extern "C" int32_t mfa_set_half_list(mfa_ctx* c, int32_t on) {
  if (!c) return MFA_ERR_ARG;
  if ((on != 0) != c->half_list) drop_graph(c);
  c->half_list = on != 0;
  const int rc = ensure_half(c);
  if (rc) c->half_list = false;
  sync_split(c);
  return rc;
}

extern "C" int32_t mfa_get_tuning(mfa_ctx* c, int32_t* tpp, int32_t* pair_queue, double* skin,
                                  int32_t* half_list) {
  if (!c) return MFA_ERR_ARG;
  if (tpp) *tpp = c->tpp;
  if (pair_queue) *pair_queue = c->pair_queue;
  if (half_list) *half_list = c->half_list ? 1 : 0;
  if (skin) *skin = c->p.skin;
  return MFA_OK;
}

extern "C" int32_t mfa_set_pair_queue(mfa_ctx* c, int32_t mode) {
  if (!c || mode < 0 || mode > 2) return MFA_ERR_ARG;
  if (mode != c->pair_queue) drop_graph(c);
  c->pair_queue = mode;
  sync_split(c);
  return MFA_OK;
}

extern "C" int32_t mfa_set_graph_steps(mfa_ctx* c, int32_t steps) {
  if (!c || steps < 0 || steps > 1024) return MFA_ERR_ARG;
  if (steps != c->graph_steps) drop_graph(c);
  c->graph_steps = steps;
  return MFA_OK;
}

// Loads replica q's state (the reference's (3,N) arrays).
static int32_t set_state_impl(mfa_ctx* c, int q, const double* r0, const double* v, const double* a,
                              const int32_t* a_type, const int32_t* mic) {
  if (!c || !r0 || !v || !a || !a_type) return MFA_ERR_ARG;
  if (q < 0 || q >= c->n_rep) return set_err(c, MFA_ERR_ARG, "replica %d out of range", q);
  const KParams& k = c->ks[(size_t)q];
  const size_t n = (size_t)k.n, o = (size_t)k.p0;
  for (size_t i = 0; i < n; ++i)
    if (a_type[i] < 1 || a_type[i] > MFA_NTYPE)
      return set_err(c, MFA_ERR_ARG, "a_type(%zu) = %d outside 1..%d", i + 1, a_type[i], MFA_NTYPE);
  HIPCHK(c, hipMemcpyAsync(c->r + 3 * o, r0, n * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->v + 3 * o, v, n * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->a + 3 * o, a, n * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->type + o, a_type, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  if (mic)
    HIPCHK(c, hipMemcpyAsync(c->mic + 3 * o, mic, n * 3 * sizeof(int32_t), hipMemcpyHostToDevice,
                             c->stream));
  else
    HIPCHK(c, hipMemsetAsync(c->mic + 3 * o, 0, n * 3 * sizeof(int32_t), c->stream));
  HIPCHK(c, hipMemsetAsync(c->flags, 0, (size_t)c->n_rep * F_NFLAGS * sizeof(int32_t), c->stream));
  if (c->dpd_vv)  // force_d starts at zero (init_params.f90:839)
    HIPCHK(c, hipMemsetAsync(c->fd + 3 * o, 0, n * 3 * sizeof(double), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->need_rebuild = true;
  if (c->ewald) {
    drop_graph(c);  // captured launches carry the charged-particle count
    return ewald_set_charges(c, a_type);
  }
  return MFA_OK;
}

static int32_t get_state_impl(mfa_ctx* c, int q, double* r0, double* v, double* a, double* force,
                              int32_t* mic) {
  if (!c) return MFA_ERR_ARG;
  if (q < 0 || q >= c->n_rep) return set_err(c, MFA_ERR_ARG, "replica %d out of range", q);
  const KParams& k = c->ks[(size_t)q];
  const size_t n = (size_t)k.n, o = (size_t)k.p0;
  if (r0)
    HIPCHK(c, hipMemcpyAsync(r0, c->r + 3 * o, n * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (v)
    HIPCHK(c, hipMemcpyAsync(v, c->v + 3 * o, n * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (a)
    HIPCHK(c, hipMemcpyAsync(a, c->a + 3 * o, n * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (force)
    HIPCHK(c, hipMemcpyAsync(force, c->f + 3 * o, n * 3 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
  if (mic)
    HIPCHK(c, hipMemcpyAsync(mic, c->mic + 3 * o, n * 3 * sizeof(int32_t), hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFA_OK;
}

extern "C" int32_t mfa_set_state(mfa_ctx* c, const double* r0, const double* v, const double* a,
                                 const int32_t* a_type, const int32_t* mic) {
  return set_state_impl(c, 0, r0, v, a, a_type, mic);
}

extern "C" int32_t mfa_get_state(mfa_ctx* c, double* r0, double* v, double* a, double* force,
                                 int32_t* mic) {
  return get_state_impl(c, 0, r0, v, a, force, mic);
}

extern "C" int32_t mfa_set_state_replica(mfa_ctx* c, int32_t replica, const double* r0,
                                         const double* v, const double* a, const int32_t* a_type,
                                         const int32_t* mic) {
  return set_state_impl(c, replica, r0, v, a, a_type, mic);
}

extern "C" int32_t mfa_get_state_replica(mfa_ctx* c, int32_t replica, double* r0, double* v,
                                         double* a, double* force, int32_t* mic) {
  return get_state_impl(c, replica, r0, v, a, force, mic);
}

extern "C" int32_t mfa_replica_obs(mfa_ctx* c, int32_t replica, mfa_obs* obs) {
  if (!c || !obs || replica < 0 || replica >= c->n_rep) return MFA_ERR_ARG;
  *obs = c->rep_obs[(size_t)replica];
  return MFA_OK;
}

extern "C" int32_t mfa_verlet_positions(mfa_ctx* c) {
  if (!c) return MFA_ERR_ARG;
  hipLaunchKernelGGL(k_verlet_positions, dim3(grid_of(c->n)), dim3(MFA_BLOCK), 0, c->stream,
                     c->d_kp, c->d_rmap, c->r, c->v, c->a, c->mic, c->r_ref, c->flags, 0, c->n_cells,
                     c->cell_count, c->type, c->cell_of, c->sidx, c->pos4, c->vel4);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFA_OK;
}

extern "C" int32_t mfa_check_skin(mfa_ctx* c, int32_t* f_skin) {
  if (!c || !f_skin) return MFA_ERR_ARG;
  if (c->need_rebuild) {
    *f_skin = 1;
    return MFA_OK;
  }
  hipLaunchKernelGGL(k_set_flag, dim3(grid_of(c->n_rep)), dim3(MFA_BLOCK), 0, c->stream, c->flags,
                     c->n_rep, (int)F_REBUILD, 0);
  hipLaunchKernelGGL(k_check_skin, dim3(grid_of(c->n)), dim3(MFA_BLOCK), 0, c->stream, c->d_kp,
                     c->d_rmap, c->r, c->r_ref, c->mic, c->flags);
  HIPCHK(c, hipGetLastError());
  int rc = check_flags(c);
  if (rc) return rc;
  *f_skin = 0;
  for (int q = 0; q < c->n_rep; ++q) *f_skin |= c->h_flags[(size_t)q * F_NFLAGS + F_REBUILD] ? 1 : 0;
  hipLaunchKernelGGL(k_set_flag, dim3(grid_of(c->n_rep)), dim3(MFA_BLOCK), 0, c->stream, c->flags,
                     c->n_rep, (int)F_REBUILD, 0);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFA_OK;
}

extern "C" int32_t mfa_binning(mfa_ctx* c) {
  if (!c) return MFA_ERR_ARG;
  c->need_rebuild = true;
  return ensure_list(c);
}

extern "C" int32_t mfa_get_cells(mfa_ctx* c, int32_t* cell_xyz) {
  if (!c || !cell_xyz) return MFA_ERR_ARG;
  int rc = ensure_list(c);
  if (rc) return rc;
  const int n0 = c->ks[0].n;  // replica 0
  std::vector<int32_t> co((size_t)n0);
  HIPCHK(c, hipMemcpy(co.data(), c->cell_of, (size_t)n0 * sizeof(int32_t), hipMemcpyDeviceToHost));
  const int nx = c->p.n_bin[0], ny = c->p.n_bin[1];
  for (int i = 0; i < n0; ++i) {
    const int cc = co[(size_t)i];
    cell_xyz[3 * (size_t)i + 0] = cc % nx;
    cell_xyz[3 * (size_t)i + 1] = (cc / nx) % ny;
    cell_xyz[3 * (size_t)i + 2] = cc / (nx * ny);
  }
  return MFA_OK;
}

extern "C" int32_t mfa_get_pairs(mfa_ctx* c, int32_t* pairs, int64_t cap, int64_t* n_pairs) {
  if (!c || !n_pairs) return MFA_ERR_ARG;
  int rc = ensure_list(c);
  if (rc) return rc;
  const int n0 = c->ks[0].n;  // replica 0 (sorted positions [0, n0))
  std::vector<int32_t> nn((size_t)n0), perm((size_t)n0);
  std::vector<unsigned> nl((size_t)n0 * c->cap);
  HIPCHK(c, hipMemcpy(nn.data(), c->nn, nn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(nl.data(), c->nl, nl.size() * sizeof(unsigned), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(perm.data(), c->perm, perm.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  int64_t m = 0;
  for (int sp = 0; sp < n0; ++sp) {
    const int cnt = nn[(size_t)sp] & 0xffff, nup = nn[(size_t)sp] >> 16;
    for (int q = 0; q < cnt; ++q) {
      const int slot = q < nup ? q : q + (c->cap - cnt);  // list_slot
      const int i = perm[(size_t)sp];
      const int j = perm[(size_t)(nl[(size_t)sp * c->cap + slot] & MFA_JMASK)];
      if (i < j) {
        if (pairs && m < cap) {
          pairs[2 * m + 0] = i + 1;
          pairs[2 * m + 1] = j + 1;
        }
        ++m;
      }
    }
  }
  *n_pairs = m;
  return MFA_OK;
}

extern "C" int32_t mfa_zero_force(mfa_ctx* c) {
  if (!c) return MFA_ERR_ARG;
  HIPCHK(c, hipMemsetAsync(c->f, 0, (size_t)c->n * 3 * sizeof(double), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFA_OK;
}

// One force-pass call with the given terms, adding into F, energies optional.
static int force_terms(mfa_ctx* c, int64_t step, int terms, double* out_terms) {
  int rc = ensure_list(c);
  if (rc) return rc;
  rc = set_step(c, step);
  if (rc) return rc;
  rc = launch_force(c, terms | TM_ACCUM | (out_terms ? TM_ENERGY : 0));
  if (rc) return rc;
  if (out_terms) {
    rc = reduce_terms(c, force_blocks(c), out_terms);
    if (rc) return rc;
  }
  return check_flags(c);
}

extern "C" int32_t mfa_fluid_fluid(mfa_ctx* c, int64_t step, double* v_ff, double* v_coul) {
  if (!c) return MFA_ERR_ARG;
  double t[MFA_NTERMS];
  int rc = force_terms(c, step, TM_PAIR, (v_ff || v_coul) ? t : nullptr);
  if (rc) return rc;
  if (v_ff) *v_ff = t[T_VFF];
  // real space plus the self term (fluid_fluid.f90:238-240)
  if (v_coul) *v_coul = c->p.coulomb ? t[T_VCOUL] + c->v_coul_self : 0.0;
  return MFA_OK;
}

extern "C" int32_t mfa_fluid_wall(mfa_ctx* c, int32_t inter_type, double* v_fw) {
  if (!c) return MFA_ERR_ARG;
  if (inter_type != 2)
    return set_err(c, MFA_ERR_ARG, "fluid_wall: only inter_type 2 (9-3 wall) is provided");
  if (c->p.symmetry != 0) {
    if (v_fw) *v_fw = 0.0;
    return MFA_OK;
  }
  double t[MFA_NTERMS];
  int rc = force_terms(c, 0, TM_WALL, v_fw ? t : nullptr);
  if (rc) return rc;
  if (v_fw) *v_fw = t[T_VFW];
  return MFA_OK;
}

extern "C" int32_t mfa_intra_molec(mfa_ctx* c, double* v_intra) {
  if (!c) return MFA_ERR_ARG;
  double t[MFA_NTERMS];
  int rc = force_terms(c, 0, TM_BOND, v_intra ? t : nullptr);
  if (rc) return rc;
  if (v_intra) *v_intra = t[T_VINTRA];
  return MFA_OK;
}

extern "C" int32_t mfa_bending(mfa_ctx* c, int32_t mode, double* v_bend) {
  if (!c) return MFA_ERR_ARG;
  if (v_bend) *v_bend = 0.0;
  if (mode == 0 || mode == 1 || !c->p.bending) return MFA_OK;  // bending.f90:12-25
  if (mode != 2) return set_err(c, MFA_ERR_ARG, "bending: mode must be 0, 1 or 2");
  double t[MFA_NTERMS];
  int rc = force_terms(c, 0, TM_BEND, v_bend ? t : nullptr);
  if (rc) return rc;
  if (v_bend) *v_bend = t[T_VBEND];
  return MFA_OK;
}

extern "C" int32_t mfa_verlet_velocities(mfa_ctx* c, double* t_kin) {
  if (!c) return MFA_ERR_ARG;
  const int g = grid_of(c->n);
  hipLaunchKernelGGL(k_verlet_velocities, dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->d_rmap,
                     c->f, c->v, c->a, c->type, c->part, t_kin ? 1 : 0, c->sidx, c->vel4,
                     c->head_f);
  HIPCHK(c, hipGetLastError());
  if (t_kin) {
    double t[MFA_NTERMS];
    int rc = reduce_terms_particles(c, t);
    if (rc) return rc;
    *t_kin = 0.5 * t[T_TKIN];
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFA_OK;
}

// ---------------------------------------------------------------------------
// verlet_velocities on arrays the caller owns, with a scratch buffer the
// caller provides (part: one row of MFA_NTERMS block partials per 256
// particles; red: MFA_NTERMS sums).  Allocates nothing; the allocation-free
// entry point (mfa_noalloc.hip) passes its file-scope device array here.
int mfa_internal_vv_scratch(const mfa_params* p, double* force, double* v, double* a,
                            const int32_t* a_type, void* stream, double* t_kin, double* part,
                            double* red) {
  KParams k = make_kparams(*p);  // p0 = 0, replica 0
  hipStream_t s = (hipStream_t)stream;
  const int g = grid_of(p->n_part);
  // every partial the sum reads ([0, g) blocks) is written by this launch,
  // and the sum overwrites its slots: nothing needs clearing
  hipLaunchKernelGGL(k_verlet_velocities_1, dim3(g), dim3(MFA_BLOCK), 0, s, k, force, v, a, a_type,
                     part, t_kin ? 1 : 0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(nullptr, MFA_ERR_HIP, "launch: %s", hipGetErrorString(e));
  if (t_kin) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(MFA_NTERMS), dim3(MFA_BLOCK), 0, s, part, 0, g, red);
    double h = 0.0;
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&h, red + T_TKIN, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_err(nullptr, MFA_ERR_HIP, "kinetic energy: %s", hipGetErrorString(e));
    *t_kin = 0.5 * h;
  }
  return MFA_OK;
}

int mfa_internal_set_error(int code, const char* msg) { return set_err(nullptr, code, "%s", msg); }

extern "C" int32_t mfa_new_dpd_fd(mfa_ctx* c, double* t_kin) {
  if (!c) return MFA_ERR_ARG;
  if (!c->dpd_vv) return set_err(c, MFA_ERR_ARG, "new_dpd_fd: DPD_VV is not set");
  int rc = ensure_list(c);
  if (rc) return rc;
  rc = launch_dpd_vv(c, t_kin ? 1 : 0, 0);
  if (rc) return rc;
  if (t_kin) {
    double t[MFA_NTERMS];
    rc = reduce_terms_particles(c, t);
    if (rc) return rc;
    *t_kin = 0.5 * t[T_TKIN];
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFA_OK;
}

static void fill_obs(const double* t, double tkin_sum, int64_t nreb, mfa_obs* obs) {
  memset(obs, 0, sizeof *obs);
  obs->v_fluid_fluid = t[T_VFF];
  obs->v_fluid_wall = t[T_VFW];
  obs->v_intra_molec = t[T_VINTRA];
  obs->v_bend = t[T_VBEND];
  obs->v_coul = t[T_VCOUL];
  obs->t_kin = 0.5 * tkin_sum;
  obs->virial = t[T_VIR];
  obs->wall_fz_bottom = t[T_FZB];
  obs->wall_fz_top = t[T_FZT];
  obs->z_moment = t[T_ZMOM];
  obs->n_rebuilds = nreb;
}

// v_coul of a full force pass: real space + self term + k space + slab dipole.
static int add_coulomb_energy(mfa_ctx* c, double* t) {
  if (!c->p.coulomb) {
    t[T_VCOUL] = 0.0;
    return MFA_OK;
  }
  double uk, udip;
  const int rc = ewald_energies(c, &uk, &udip);
  if (rc) return rc;
  t[T_VCOUL] = t[T_VCOUL] + c->v_coul_self + uk + udip;
  return MFA_OK;
}

extern "C" int32_t mfa_ewald_k(mfa_ctx* c, int32_t mode, double* v_coul) {
  if (!c) return MFA_ERR_ARG;
  if (v_coul) *v_coul = 0.0;
  if (mode == 0 || !ewald_active(c)) return MFA_OK;  // the k grid is set up by mfa_create
  if (mode != 1) return set_err(c, MFA_ERR_ARG, "ewald_k: mode must be 0 or 1");
  int rc = ensure_list(c);
  if (rc) return rc;
  rc = launch_ewald(c, 0);
  if (rc) return rc;
  double uk, udip;
  rc = ewald_energies(c, &uk, &udip);
  if (rc) return rc;
  if (v_coul) *v_coul = uk;
  return check_flags(c);
}

extern "C" int32_t mfa_dipolar_correction(mfa_ctx* c, double* v_coul) {
  if (!c) return MFA_ERR_ARG;
  if (v_coul) *v_coul = 0.0;
  if (!ewald_active(c) || c->p.symmetry != 0) return MFA_OK;
  const EwaldK& e = c->ew.e;
  hipLaunchKernelGGL(k_ewald_mz, dim3(1), dim3(MFA_BLOCK), 0, c->stream, e, c->r, c->d_qidx,
                     c->d_qval, c->d_mz);
  hipLaunchKernelGGL(k_ewald_apply, dim3((3 * e.nq + MFA_BLOCK - 1) / MFA_BLOCK), dim3(MFA_BLOCK), 0,
                     c->stream, e, c->d_qidx, c->d_qval, c->d_fk_part, c->ew.n_kgroups, c->d_mz,
                     c->ew.inv_V, 0, 1, c->f);
  HIPCHK(c, hipGetLastError());
  double uk, udip;
  int rc = ewald_energies(c, &uk, &udip);
  if (rc) return rc;
  if (v_coul) *v_coul = udip;
  return check_flags(c);
}

extern "C" int32_t mfa_energies(mfa_ctx* c, int64_t step, mfa_obs* obs) {
  if (!c || !obs) return MFA_ERR_ARG;
  int rc = ensure_list(c);
  if (rc) return rc;
  rc = set_step(c, step);
  if (rc) return rc;
  // recompute forces into F (the pass overwrites F), then reduce
  rc = launch_force(c, TM_ALL | TM_ENERGY);
  if (rc) return rc;
  rc = launch_ewald(c, c->p.symmetry == 0);
  if (rc) return rc;
  double t[MFA_NTERMS];
  rc = reduce_terms(c, force_blocks(c), t);
  if (rc) return rc;
  rc = add_coulomb_energy(c, t);
  if (rc) return rc;
  // kinetic energy of the current velocities
  const int n0 = c->ks[0].n;  // replica 0
  std::vector<double> vv((size_t)n0 * 3);
  std::vector<int32_t> ty((size_t)n0);
  HIPCHK(c, hipMemcpy(vv.data(), c->v, vv.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(ty.data(), c->type, ty.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  double tk = 0.0;
  for (int i = 0; i < n0; ++i) {
    const double* q = &vv[3 * (size_t)i];
    tk += c->p.mass_type[ty[(size_t)i] - 1] * ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]);
  }
  fill_obs(t, tk, 0, obs);
  return check_flags(c);
}

extern "C" int32_t mfa_run_steps(mfa_ctx* c, int64_t first_step, int32_t n, mfa_obs* obs) {
  if (!c || n < 0) return MFA_ERR_ARG;
  if (n == 0) return MFA_OK;
  int rc = ensure_list(c);
  if (rc) return rc;
  hipLaunchKernelGGL(k_set_flag, dim3(grid_of(c->n_rep)), dim3(MFA_BLOCK), 0, c->stream, c->flags,
                     c->n_rep, (int)F_NREBUILD, 0);
  // with the integration inside the force pass, each step's rebuild moves
  // the step counter on before that step's force pass
  const bool fused = fused_ok(c);
  rc = set_step(c, fused ? first_step - 1 : first_step);
  if (rc) return rc;
  const int g = grid_of(c->n);
  // first half of the first step (verlet_positions + check_skin)
  hipLaunchKernelGGL(k_verlet_positions, dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->d_rmap,
                     c->r, c->v, c->a, c->mic, c->r_ref, c->flags, 1, c->n_cells, c->cell_count,
                     c->type, c->cell_of, c->sidx, c->pos4, c->vel4);
  HIPCHK(c, hipGetLastError());
  // steps 1 .. n-1: captured blocks of graph_steps steps, then single steps
  int middle = n - 1;
  // a captured block of fused steps must leave the buffers where it found
  // them: an even number of steps
  const bool graphs = c->graph_steps > 0 && !c->timing && !(fused && (c->graph_steps & 1));
  if (graphs && middle >= c->graph_steps) {
    rc = ensure_graph(c);
    if (rc) return rc;
    while (middle >= c->graph_steps) {
      HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
      middle -= c->graph_steps;
    }
  }
  for (; middle > 0; --middle) {
    rc = enqueue_middle_step(c);
    if (rc) return rc;
  }
  // last step: force pass with the energies, verlet_velocities alone
  rc = launch_rebuild(c, nullptr, fused ? 1 : 0);
  if (rc) return rc;
  rc = launch_force(c, TM_ALL | (obs ? TM_ENERGY : 0));
  if (rc) return rc;
  rc = launch_ewald(c, c->p.symmetry == 0);
  if (rc) return rc;
  const size_t nt = (size_t)c->n_rep * MFA_NTERMS;
  std::vector<double> t(nt), tk(nt);
  if (obs) {
    hipLaunchKernelGGL(k_reduce_replicas, dim3(MFA_NTERMS, c->n_rep), dim3(MFA_BLOCK), 0, c->stream,
                       c->part, c->d_kp, MFA_BLOCK / c->tpp, c->red);
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->red, nt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < nt; ++q) t[q] = c->h_red[q];
    rc = add_coulomb_energy(c, t.data());  // replica 0 (charged systems run alone)
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_verlet_velocities, dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->d_rmap,
                     c->f, c->v, c->a, c->type, c->part, (obs && !c->dpd_vv) ? 1 : 0, c->sidx,
                     c->vel4, c->head_f);
  HIPCHK(c, hipGetLastError());
  if (c->dpd_vv) {  // md_main.f90:149-155; the kinetic energy comes from new_dpd_fd
    rc = launch_dpd_vv(c, obs ? 1 : 0, 0);
    if (rc) return rc;
  }
  if (obs) {
    hipLaunchKernelGGL(k_reduce_replicas, dim3(MFA_NTERMS, c->n_rep), dim3(MFA_BLOCK), 0, c->stream,
                       c->part, c->d_kp, MFA_BLOCK, c->red);
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->red, nt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < nt; ++q) tk[q] = c->h_red[q];
  }
  rc = check_flags(c);
  if (rc) return rc;
  if (obs) {
    for (int q = 0; q < c->n_rep; ++q)
      fill_obs(&t[(size_t)q * MFA_NTERMS], tk[(size_t)q * MFA_NTERMS + T_TKIN],
               c->h_flags[(size_t)q * F_NFLAGS + F_NREBUILD], &c->rep_obs[(size_t)q]);
    *obs = c->rep_obs[0];
  }
  return MFA_OK;
}

// ---------------------------------------------------------------------------
// Observation stage (observation.f90:17-307 with PROFILES; functions.f90
// histo).  Per particle: the momentum partials m v and the slab counts of z
// (brush, melt, particle 4).  Per chain: the chain unfolded along its bonds
// (one image correction per monomer and axis, observation.f90:163-199) and
// its Rg^2 about its centre (:205-233).  A one-block kernel then adds the
// observation to the device sums in a fixed order.
__global__ __launch_bounds__(MFA_BLOCK) void k_obs_particles(
    const KParams* __restrict__ kb, const int* __restrict__ rmap, const double* __restrict__ r,
    const double* __restrict__ v, const int* __restrict__ type, unsigned* __restrict__ hist,
    int hist_dim, double inv_z_step, double* __restrict__ part) {
  MFA_REPLICA(blockIdx.x * blockDim.x);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double vals[MFA_NTERMS];
  for (int t = 0; t < MFA_NTERMS; ++t) vals[t] = 0.0;
  if (i - k.p0 < k.n) {
    const double m = k.mass_type[type[i] - 1];
    for (int d = 0; d < 3; ++d) vals[d] = m * v[3 * (size_t)i + d];
    if (hist_dim > 0) {
      const int il = i - k.p0;
      const int g = il < k.n_brush ? 0 : (il < k.part_init_e ? 1 : 2);
      int zi = (int)(r[3 * (size_t)i + 2] * inv_z_step) + 1;  // functions.f90:148-156
      zi = zi < 1 ? 1 : (zi > hist_dim ? hist_dim : zi);
      atomicAdd(&hist[(size_t)g * hist_dim + (zi - 1)], 1u);
    }
  }
  block_partials(vals, part);
}

// One chain's unfolded monomer m: positions from the chain start, each
// corrected by one box length when it sits more than half a box from the
// (corrected) previous one along an axis.
__global__ __launch_bounds__(MFA_BLOCK) void k_obs_chains(const KParams* __restrict__ kb,
                                                          const double* __restrict__ r,
                                                          double* __restrict__ chain_rg) {
  const KParams& k = kb[0];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int n_chain_d = (k.part_init_e - k.n_brush) / (k.n_mon_d > 0 ? k.n_mon_d : 1);
  if (c >= k.n_chain + n_chain_d) return;
  const bool brush = c < k.n_chain;
  const int len = brush ? k.n_mon : k.n_mon_d;
  const int base = brush ? c * k.n_mon : k.n_brush + (c - k.n_chain) * k.n_mon_d;
  const bool unfold = brush ? k.symmetry == 0 : true;
  double half[3];
  for (int d = 0; d < 3; ++d) half[d] = k.L[d] / 2.;  // init_config.f90:22
  double cm[3] = {0., 0., 0.}, vd[3] = {0., 0., 0.};
  for (int pass = 0; pass < 2; ++pass) {
    double prev[3] = {0., 0., 0.};
    for (int mm = 0; mm < len; ++mm) {
      const int i = base + mm;
      double x[3];
      for (int d = 0; d < 3; ++d) {
        x[d] = r[3 * (size_t)i + d];
        // mod(i_part-1, n_mon) on the global index, as the reference does
        const bool first = brush ? (i % k.n_mon) == 0 : (i % k.n_mon_d) == 0;
        if (unfold && !first) {
          if (x[d] - prev[d] > half[d]) x[d] = x[d] - k.L[d];
          else if (prev[d] - x[d] > half[d]) x[d] = x[d] + k.L[d];
        }
        prev[d] = x[d];
        if (pass == 0) {
          cm[d] = cm[d] + x[d];
        } else {
          const double y = x[d] - cm[d];
          vd[d] = vd[d] + y * y;
        }
      }
    }
    if (pass == 0)
      for (int d = 0; d < 3; ++d) cm[d] = brush ? cm[d] * (1. / (double)k.n_mon) : cm[d] / (double)k.n_mon_d;
  }
  for (int d = 0; d < 3; ++d)
    chain_rg[3 * (size_t)c + d] = brush ? vd[d] * (1. / (double)k.n_mon) : vd[d] / (double)k.n_mon_d;
}

// Adds one observation to the device sums: obs[0..2] tot_P, [3..11]
// r_g2_mean, [12..13] r_par_per2 (observation.f90:232-291).  Fixed-order sums.
__global__ __launch_bounds__(MFA_BLOCK) void k_obs_finish(const KParams* __restrict__ kb,
                                                          const double* __restrict__ part, int nblocks,
                                                          const double* __restrict__ chain_rg,
                                                          int with_rg, double* __restrict__ obs) {
  const KParams& k = kb[0];
  __shared__ double sh[MFA_BLOCK][3];
  const int t = threadIdx.x;
  const int n_chain_d = (k.part_init_e - k.n_brush) / (k.n_mon_d > 0 ? k.n_mon_d : 1);
  double res[3][3];  // momentum, brush Rg^2 sum, melt Rg^2 sum
  for (int q = 0; q < 3; ++q) {
    double a[3] = {0., 0., 0.};
    if (q == 0) {
      for (int b = t; b < nblocks; b += MFA_BLOCK)
        for (int d = 0; d < 3; ++d) a[d] += part[(size_t)b * MFA_NTERMS + d];
    } else if (with_rg) {
      const int c0 = q == 1 ? 0 : k.n_chain, c1 = q == 1 ? k.n_chain : k.n_chain + n_chain_d;
      for (int c = c0 + t; c < c1; c += MFA_BLOCK)
        for (int d = 0; d < 3; ++d) a[d] += chain_rg[3 * (size_t)c + d];
    }
    for (int d = 0; d < 3; ++d) sh[t][d] = a[d];
    __syncthreads();
    for (int o = MFA_BLOCK / 2; o > 0; o >>= 1) {
      if (t < o)
        for (int d = 0; d < 3; ++d) sh[t][d] += sh[t + o][d];
      __syncthreads();
    }
    for (int d = 0; d < 3; ++d) res[q][d] = sh[0][d];
    __syncthreads();
  }
  if (t != 0) return;
  for (int d = 0; d < 3; ++d) obs[d] = obs[d] + res[0][d];
  if (!with_rg) return;
  double rg[2][3];
  for (int d = 0; d < 3; ++d) {
    rg[0][d] = res[1][d] * (1. / (double)k.n_chain);
    rg[1][d] = res[2][d] / (double)n_chain_d;  // 0/0 without melt chains, as in the reference
  }
  for (int d = 0; d < 3; ++d) {
    obs[3 + d] = obs[3 + d] + rg[0][d];
    obs[6 + d] = obs[6 + d] + rg[1][d];
    obs[9 + d] = obs[9 + d] + (rg[0][d] + rg[1][d]) / 2.0;
  }
  for (int g = 0; g < 2; ++g) {
    const double sum = (rg[g][0] + rg[g][1]) + rg[g][2];
    obs[12 + g] = obs[12 + g] + (rg[g][2] - 0.5 * (rg[g][0] + rg[g][1])) / sum;
  }
}

extern "C" int32_t mfa_obs_init(mfa_ctx* c, int32_t hist_dim) {
  if (!c) return MFA_ERR_ARG;
  if (c->n_rep != 1) return set_err(c, MFA_ERR_ARG, "observation: single-system contexts only");
  if (hist_dim < 0 || hist_dim > (1 << 20)) return set_err(c, MFA_ERR_ARG, "observation: bad hist_dim %d", hist_dim);
  const mfa_params& p = c->p;
  if (!c->head_f) HIPCHK(c, hipMalloc(&c->head_f, 3 * (size_t)(p.n_chain > 0 ? p.n_chain : 1) * sizeof(double)));
  if (!c->d_obs) HIPCHK(c, hipMalloc(&c->d_obs, 16 * sizeof(double)));
  if (!c->d_chain_rg)
    HIPCHK(c, hipMalloc(&c->d_chain_rg, 3 * (size_t)(p.n_chain + p.n_chain_d + 1) * sizeof(double)));
  if (c->d_hist && hist_dim != c->hist_dim) {
    hipFree(c->d_hist);
    c->d_hist = nullptr;
  }
  if (hist_dim > 0 && !c->d_hist) HIPCHK(c, hipMalloc(&c->d_hist, 3 * (size_t)hist_dim * sizeof(unsigned)));
  c->hist_dim = hist_dim;
  HIPCHK(c, hipMemsetAsync(c->head_f, 0, 3 * (size_t)(p.n_chain > 0 ? p.n_chain : 1) * sizeof(double), c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_obs, 0, 16 * sizeof(double), c->stream));
  if (hist_dim > 0) HIPCHK(c, hipMemsetAsync(c->d_hist, 0, 3 * (size_t)hist_dim * sizeof(unsigned), c->stream));
  memset(&c->obs_host, 0, sizeof c->obs_host);
  c->obs_host.hist_dim = hist_dim;
  c->obs_on = true;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFA_OK;
}

extern "C" int32_t mfa_observe(mfa_ctx* c, const mfa_obs* e, int32_t coul_in_total) {
  if (!c || !e) return MFA_ERR_ARG;
  if (!c->obs_on) return set_err(c, MFA_ERR_ARG, "observation: call mfa_obs_init first");
  const mfa_params& p = c->p;
  mfa_obs_sums* o = &c->obs_host;
  // the energy totals of verlet_velocities.f90:183-196 and the sums of
  // observation.f90:129-159 (no wall atoms: their terms are zero)
  const double v_ff = e->v_fluid_fluid, v_fw = e->v_fluid_wall, v_in = e->v_intra_molec;
  const double t_fluid = e->t_kin;
  o->count_obs += 1;
  double v_total = ((0. + v_fw) + v_ff) + v_in;
  const double t_total = 0. + t_fluid;
  const double e_total = v_total + t_total;
  if (coul_in_total) v_total = v_total + e->v_coul;
  const double vf = (v_ff + v_in) + v_fw;
  o->v_fluid_fluid_1 += vf;
  o->t_fluid_1 += t_fluid;
  o->e_fluid_fluid_1 += v_ff + t_fluid;
  o->v_fluid_fluid_2 += vf * vf;
  o->t_fluid_2 += t_fluid * t_fluid;
  o->e_fluid_fluid_2 += (v_ff + t_fluid) * (v_ff + t_fluid);
  o->v_total_1 += v_total;
  o->t_total_1 += t_total;
  o->e_total_1 += e_total;
  o->v_total_2 += v_total * v_total;
  o->t_total_2 += t_total * t_total;
  o->e_total_2 += e_total * e_total;
  o->v_tot_intra_1 += v_in;
  o->v_tot_intra_2 += v_in * v_in;
  const int g = grid_of(c->n);
  const double inv_z_step = 1. / (p.boundary[2] / (double)(c->hist_dim > 0 ? c->hist_dim : 1));
  hipLaunchKernelGGL(k_obs_particles, dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->d_rmap, c->r,
                     c->v, c->type, c->d_hist, c->hist_dim, inv_z_step, c->part);
  const int with_rg = p.system == 3 ? 1 : 0;  // observation.f90:201 (SYSTEM 0 or 3)
  const int nch = p.n_chain + p.n_chain_d;
  if (with_rg && nch > 0)
    hipLaunchKernelGGL(k_obs_chains, dim3((nch + MFA_BLOCK - 1) / MFA_BLOCK), dim3(MFA_BLOCK), 0,
                       c->stream, c->d_kp, c->r, c->d_chain_rg);
  hipLaunchKernelGGL(k_obs_finish, dim3(1), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->part, g,
                     c->d_chain_rg, with_rg && nch > 0 ? 1 : 0, c->d_obs);
  HIPCHK(c, hipGetLastError());
  if (c->hist_dim > 0) o->count_hist += 1;
  return MFA_OK;
}

extern "C" int32_t mfa_obs_read(mfa_ctx* c, mfa_obs_sums* out, double* dens_z, double* f_on_heads) {
  if (!c) return MFA_ERR_ARG;
  if (!c->obs_on) return set_err(c, MFA_ERR_ARG, "observation: call mfa_obs_init first");
  const mfa_params& p = c->p;
  double dv[16];
  HIPCHK(c, hipMemcpyAsync(dv, c->d_obs, 16 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  std::vector<unsigned> h;
  if (dens_z && c->hist_dim > 0) {
    h.resize(3 * (size_t)c->hist_dim);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_hist, h.size() * sizeof(unsigned), hipMemcpyDeviceToHost,
                             c->stream));
  }
  if (f_on_heads && p.n_chain > 0)
    HIPCHK(c, hipMemcpyAsync(f_on_heads, c->head_f, 3 * (size_t)p.n_chain * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (out) {
    *out = c->obs_host;
    for (int d = 0; d < 3; ++d) out->tot_p[d] = dv[d];
    for (int q = 0; q < 3; ++q)
      for (int d = 0; d < 3; ++d) out->r_g2_mean[q][d] = dv[3 + 3 * q + d];
    out->r_par_per2[0] = dv[12];
    out->r_par_per2[1] = dv[13];
  }
  for (size_t q = 0; q < h.size(); ++q) dens_z[q] = (double)h[q];
  return MFA_OK;
}

extern "C" int32_t mfa_autotune(mfa_ctx* c, int64_t first_step, int32_t n_steps, int32_t* best_tpp,
                                double* ms_per_step) {
  if (!c || n_steps < 1) return MFA_ERR_ARG;
  const int cand[5] = {1, 2, 4, 8, 16};
  double best = 1e300;
  int best_t = c->tpp, best_q = c->pair_queue;
  bool best_h = c->half_list;
  // How the mixed pass walks the list, tried with every width: directly; the
  // skin part through the pair queue and the queue for everything (compiled
  // for systems without extra terms; the latter for DPD_CUT_OFF only, where
  // few candidates interact); the half-list form (4 or more lanes), also
  // with the queue for DPD_CUT_OFF.
  struct Variant {
    int queue;
    bool half;
  };
  std::vector<Variant> vars = {{0, false}};
  const bool mixed = c->precision == MFA_MIXED;
  if (mixed) {
    const bool qok = !c->extra_terms && !c->lgv_terms, dc = c->p.dpd_cut_off_on != 0;
    if (qok) vars.push_back({2, false});
    if (qok && dc) vars.push_back({1, false});
    vars.push_back({0, true});
    if (qok && dc) vars.push_back({1, true});
  }
  int64_t step = first_step;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  const bool timing = c->timing;
  c->timing = false;
  int rc = MFA_OK;
  // two sweeps over the candidates, each candidate keeping its faster trial:
  // the state moves on between trials, and the rebuilds a trial meets vary
  std::vector<double> best_of(vars.size() * 5, 1e300);
  for (int sweep = 0; sweep < 2 && rc == MFA_OK; ++sweep)
  for (size_t vi = 0; vi < vars.size() && rc == MFA_OK; ++vi) {
    for (int q = 0; q < 5 && rc == MFA_OK; ++q) {
      if (vars[vi].half && cand[q] < 4) continue;  // the half-list passes need 4 or more lanes
      c->tpp = cand[q];
      c->pair_queue = vars[vi].queue;
      if (mixed) rc = mfa_set_half_list(c, vars[vi].half ? 1 : 0);  // also sets the row layout
      if (rc) break;
      sync_split(c);
      drop_graph(c);
      // the candidate's first steps (untimed; they capture its step block),
      // then n_steps whole steps as the run makes them (the force pass as the
      // step runs it, and the rebuilds with their row layout), timed
      rc = mfa_run_steps(c, step, c->graph_steps + 1, nullptr);
      step += c->graph_steps + 1;
      if (rc) break;
      HIPCHK(c, hipEventRecord(e0, c->stream));
      rc = mfa_run_steps(c, step, n_steps, nullptr);
      step += n_steps;
      if (rc) break;
      HIPCHK(c, hipEventRecord(e1, c->stream));
      HIPCHK(c, hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
      double& bo = best_of[vi * 5 + (size_t)q];
      bo = ms < bo ? ms : bo;
      if (sweep == 1 && bo < best) {
        best = bo;
        best_t = cand[q];
        best_q = vars[vi].queue;
        best_h = vars[vi].half;
      }
    }
  }
  c->timing = timing;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  c->tpp = best_t;
  c->pair_queue = best_q;
  if (mixed) {
    const int rh = mfa_set_half_list(c, best_h ? 1 : 0);
    if (!rc) rc = rh;
  }
  sync_split(c);
  drop_graph(c);
  if (rc) return rc;
  if (best_tpp) *best_tpp = best_t;
  if (ms_per_step) *ms_per_step = best / (double)n_steps;
  return check_flags(c);
}

extern "C" int32_t mfa_set_skin(mfa_ctx* c, double skin) {
  if (!c) return MFA_ERR_ARG;
  std::vector<mfa_params> ps = c->ps;
  int cap = c->cap_fixed ? c->cap : 0;
  for (size_t q = 0; q < ps.size(); ++q) {
    if (mfa_params_set_skin(&ps[q], skin) != MFA_OK)
      return set_err(c, MFA_ERR_ARG, "replica %zu: %s", q, mfa_last_error(nullptr));
    if (!c->cap_fixed) {
      const int cq = list_cap_of(ps[q]);
      cap = cq > cap ? cq : cap;
    }
  }
  if (cap > MFA_CAP_MAX)  // row counts are packed in 16 bits
    return set_err(c, MFA_ERR_ARG, "list capacity %d beyond %d at skin %g", cap, MFA_CAP_MAX, skin);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (cap > c->cap) {  // a longer list: grow its buffer (its contents are rebuilt)
    unsigned* nl = nullptr;
    const size_t bytes = (size_t)c->n * (size_t)cap * sizeof(unsigned);
    if (hipMalloc((void**)&nl, bytes) != hipSuccess)
      return set_err(c, MFA_ERR_NOMEM, "hipMalloc of %zu bytes (list) failed", bytes);
    hipFree(c->nl);
    c->nl = nl;
    c->cap = cap;
    const int rc = ensure_half(c);  // the half-list buffers follow the capacity
    if (rc) return rc;
  }
  for (size_t q = 0; q < ps.size(); ++q) kparams_skin(c->ks[q], ps[q]);
  HIPCHK(c, hipMemcpy(c->d_kp, c->ks.data(), c->ks.size() * sizeof(KParams), hipMemcpyHostToDevice));
  c->ps = ps;
  c->p = ps[0];
  c->need_rebuild = true;  // a new list (and skin reference) before the next force pass
  drop_graph(c);
  return MFA_OK;
}

// Times n_steps whole steps at each candidate skin that fits the cells
// (after graph_steps + 1 untimed ones), from *step on (advanced past every
// step run); the fastest skin and its time for the n_steps (ms, 1e300 if no
// candidate fits).  Leaves the context at the last candidate tried.
static int skin_trials(mfa_ctx* c, int64_t* step_io, int32_t n_steps, int32_t n_cand,
                       const double* skins, double* best_s, double* best_ms) {
  double best = 1e300;
  *best_s = c->p.skin;
  int64_t step = *step_io;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  int rc = MFA_OK;
  for (int q = 0; q < n_cand && rc == MFA_OK; ++q) {
    if (mfa_set_skin(c, skins[q]) != MFA_OK) continue;  // too wide for the cells: skipped
    // the candidate's first steps (untimed; they capture its step block)
    rc = mfa_run_steps(c, step, c->graph_steps + 1, nullptr);
    step += c->graph_steps + 1;
    if (rc) break;
    HIPCHK(c, hipEventRecord(e0, c->stream));
    rc = mfa_run_steps(c, step, n_steps, nullptr);
    step += n_steps;
    if (rc) break;
    HIPCHK(c, hipEventRecord(e1, c->stream));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      *best_s = skins[q];
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *step_io = step;
  *best_ms = best;
  return rc;
}

extern "C" int32_t mfa_autotune_skin(mfa_ctx* c, int64_t first_step, int32_t n_steps,
                                     int32_t n_cand, const double* skins, double* best_skin,
                                     double* ms_per_step) {
  if (!c || n_steps < 1 || n_cand < 1 || n_cand > 64 || !skins) return MFA_ERR_ARG;
  int64_t step = first_step;
  double best = 1e300, best_s = c->p.skin;
  int rc = skin_trials(c, &step, n_steps, n_cand, skins, &best_s, &best);
  if (rc) return rc;
  if (best > 1e299) return set_err(c, MFA_ERR_ARG, "no candidate skin fits the cells");
  rc = mfa_set_skin(c, best_s);
  if (rc) return rc;
  if (best_skin) *best_skin = best_s;
  if (ms_per_step) *ms_per_step = best / (double)n_steps;
  return check_flags(c);
}

extern "C" int32_t mfa_set_cells(mfa_ctx* c, double width) {
  if (!c) return MFA_ERR_ARG;
  // Replica batches keep the reference grid: the cell-size autotuner on the
  // 256-replica batch ended in a GPU memory fault whose cause is not found.
  if (c->n_rep > 1)
    return set_err(c, MFA_ERR_ARG, "another cell grid: single-system contexts only");
  std::vector<mfa_params> ps = c->ps;
  for (size_t q = 0; q < ps.size(); ++q)
    if (mfa_params_set_cells(&ps[q], width) != MFA_OK)
      return set_err(c, MFA_ERR_ARG, "replica %zu: %s", q, mfa_last_error(nullptr));
  // the replicas' cell ranges and the common cell capacity, as at creation
  std::vector<KParams> ks = c->ks;
  int cell_cap = 0;
  int64_t ctot = 0, np = 0;
  for (size_t q = 0; q < ps.size(); ++q) {
    const mfa_params& p = ps[q];
    const double dens = (double)p.n_part / (p.boundary[0] * p.boundary[1] * p.boundary[2]);
    const double occ = dens * p.r_bin[0] * p.r_bin[1] * p.r_bin[2];
    const int cc = (((int)(2.0 * occ + 32.0)) + 31) / 32 * 32;
    cell_cap = cc > cell_cap ? cc : cell_cap;
    KParams k = make_kparams(p);
    k.p0 = ks[q].p0;
    k.fg0 = ks[q].fg0;
    k.rep = ks[q].rep;
    k.c0 = (int)ctot;
    ks[q] = k;
    ctot += (int64_t)p.n_bin[0] * p.n_bin[1] * p.n_bin[2];
    np += p.n_part;
  }
  if (cell_cap > LB_MAX_CELL)
    return set_err(c, MFA_ERR_ARG, "cells of width %g too crowded for the list build (capacity %d)",
                   width, cell_cap);
  if (ctot > 0x7fffffff) return set_err(c, MFA_ERR_ARG, "too many cells");
  for (KParams& k : ks) k.cell_cap = cell_cap;
  const int old_cap = c->cell_cap;
  c->cell_cap = cell_cap;
  if (list_lds_bytes(c) > c->lds_limit) {
    const size_t b = list_lds_bytes(c);
    c->cell_cap = old_cap;
    return set_err(c, MFA_ERR_ARG, "cells of width %g: the list build needs %zu bytes of LDS", width, b);
  }
  c->cell_cap = old_cap;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t nc = (size_t)ctot;
  int32_t *count = nullptr, *slots = nullptr, *start = nullptr, *crep_d = nullptr;
  if (hipMalloc((void**)&count, nc * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void**)&slots, nc * (size_t)cell_cap * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void**)&start, (nc + 1) * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void**)&crep_d, nc * sizeof(int32_t)) != hipSuccess) {
    hipFree(count);
    hipFree(slots);
    hipFree(start);
    hipFree(crep_d);
    return set_err(c, MFA_ERR_NOMEM, "hipMalloc of the cell arrays (%zu cells) failed", nc);
  }
  std::vector<int> crep(nc, 0);
  for (size_t q = 0; q < ks.size(); ++q)
    for (int cc = 0; cc < ks[q].n_cells; ++cc) crep[(size_t)(ks[q].c0 + cc)] = (int)q;
  if (hipMemcpy(crep_d, crep.data(), nc * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_kp, ks.data(), ks.size() * sizeof(KParams), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(count, 0, nc * sizeof(int32_t)) != hipSuccess) {
    hipFree(count);
    hipFree(slots);
    hipFree(start);
    hipFree(crep_d);
    return set_err(c, MFA_ERR_HIP, "upload of the cell tables failed");
  }
  hipFree(c->cell_count);
  hipFree(c->cell_slots);
  hipFree(c->cell_start);
  hipFree(c->d_cell_rep);
  c->cell_count = count;
  c->cell_slots = slots;
  c->cell_start = start;
  c->d_cell_rep = crep_d;
  c->ks = ks;
  c->ps = ps;
  c->p = ps[0];
  c->n_cells = (int)ctot;
  c->cell_cap = cell_cap;
  c->lb_octants = ctot > 0 && np >= (int64_t)LB_MIN_OCTANT * ctot;
  c->need_rebuild = true;  // new cells (and a new list) before the next force pass
  drop_graph(c);
  return MFA_OK;
}

extern "C" int32_t mfa_autotune_cells(mfa_ctx* c, int64_t first_step, int32_t n_steps,
                                      int32_t n_width, const double* widths, int32_t n_skin,
                                      const double* skins, double* best_width, double* best_skin,
                                      double* ms_per_step, int64_t* next_step) {
  if (!c || n_steps < 1 || n_width < 1 || n_width > 16 || !widths || n_skin < 1 || n_skin > 64 ||
      !skins)
    return MFA_ERR_ARG;
  int64_t step = first_step;
  double best = 1e300, best_w = 0.0, best_s = c->p.skin;
  int rc = MFA_OK;
  double smin = skins[0];
  for (int q = 1; q < n_skin; ++q) smin = skins[q] < smin ? skins[q] : smin;
  for (int q = 0; q < n_width && rc == MFA_OK; ++q) {
    // the skin of the last trial may not fit these cells: retry with the
    // smallest candidate skin; a width that still does not fit is skipped
    if (mfa_set_cells(c, widths[q]) != MFA_OK &&
        (mfa_set_skin(c, smin) != MFA_OK || mfa_set_cells(c, widths[q]) != MFA_OK))
      continue;
    double s = 0.0, ms = 1e300;
    rc = skin_trials(c, &step, n_steps, n_skin, skins, &s, &ms);
    if (!rc && ms < best) {
      best = ms;
      best_w = widths[q];
      best_s = s;
    }
  }
  if (rc) return rc;
  if (best > 1e299) return set_err(c, MFA_ERR_ARG, "no candidate cell width and skin fit");
  // best_s fits best_w's cells and so any wider ones: set the skin first
  if (mfa_set_skin(c, best_s) != MFA_OK) {
    if ((rc = mfa_set_skin(c, smin))) return rc;
  }
  if ((rc = mfa_set_cells(c, best_w)) || (rc = mfa_set_skin(c, best_s))) return rc;
  if (best_width) *best_width = best_w;
  if (best_skin) *best_skin = best_s;
  if (ms_per_step) *ms_per_step = best / (double)n_steps;
  if (next_step) *next_step = step;
  return check_flags(c);
}

// Average duration of `reps` launches of `launch` on the context's stream,
// captured in one graph when graph steps are on, bracketed by two HIP events
// (a first untimed round warms up).
static int time_launches(mfa_ctx* c, int reps, int (*launch)(mfa_ctx*), double* ms_avg) {
  int rc = MFA_OK;
  hipGraphExec_t ge = nullptr;
  if (c->graph_steps > 0) {  // back-to-back launches captured in one graph
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    for (int q = 0; q < reps && rc == MFA_OK; ++q) rc = launch(c);
    c->capturing = false;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return set_err(c, MFA_ERR_HIP, "stream capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return set_err(c, MFA_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
  }
  const bool timing = c->timing;
  c->timing = false;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  for (int pass = 0; pass < 2 && rc == MFA_OK; ++pass) {  // pass 0 warms up
    if (pass == 1) HIPCHK(c, hipEventRecord(e0, c->stream));
    if (ge) {
      HIPCHK(c, hipGraphLaunch(ge, c->stream));
    } else {
      for (int q = 0; q < reps && rc == MFA_OK; ++q) rc = launch(c);
    }
  }
  c->timing = timing;
  HIPCHK(c, hipEventRecord(e1, c->stream));
  HIPCHK(c, hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (ge) hipGraphExecDestroy(ge);
  if (rc) return rc;
  *ms_avg = (double)ms / (double)reps;
  return MFA_OK;
}

static int launch_force_all(mfa_ctx* c) { return launch_force(c, TM_ALL); }

extern "C" int32_t mfa_time_force(mfa_ctx* c, int32_t reps, double* ms_avg) {
  if (!c || !ms_avg || reps < 1 || reps > 4096) return MFA_ERR_ARG;
  int rc = ensure_list(c);
  if (rc) return rc;
  rc = time_launches(c, reps, launch_force_all, ms_avg);
  return rc ? rc : check_flags(c);
}

// The force pass as the step runs it (see the header): in the fused form
// its kernel also integrates, and is timed with its writes to v, mic, the
// flags and the cell counts sent to copies, so the state stays as it was.
static int launch_force_fused(mfa_ctx* c) { return launch_force(c, TM_ALL | TM_FUSE); }

extern "C" int32_t mfa_time_force_step(mfa_ctx* c, int32_t reps, double* ms_avg) {
  if (!c || !ms_avg || reps < 1 || reps > 4096) return MFA_ERR_ARG;
  if (!fused_ok(c)) return mfa_time_force(c, reps, ms_avg);
  int rc = ensure_list(c);
  if (rc) return rc;
  const size_t n3 = 3 * (size_t)c->n, nf = (size_t)c->n_rep * F_NFLAGS;
  double* pv = nullptr;
  int32_t *pm = nullptr, *pf = nullptr, *pc = nullptr;
  if (hipMalloc(&pv, n3 * sizeof(double)) != hipSuccess ||
      hipMalloc(&pm, n3 * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&pf, nf * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&pc, (size_t)c->n_cells * sizeof(int32_t)) != hipSuccess) {
    hipFree(pv);
    hipFree(pm);
    hipFree(pf);
    hipFree(pc);
    return set_err(c, MFA_ERR_NOMEM, "hipMalloc of the timing copies failed");
  }
  if (hipMemcpyAsync(pf, c->flags, nf * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream) !=
      hipSuccess)
    rc = set_err(c, MFA_ERR_HIP, "flag copy failed");
  if (!rc) {
    c->probe_v = pv;
    c->probe_mic = pm;
    c->probe_flags = pf;
    c->probe_count = pc;
    rc = time_launches(c, reps, launch_force_fused, ms_avg);
    c->probe_v = nullptr;
    c->probe_mic = nullptr;
    c->probe_flags = nullptr;
    c->probe_count = nullptr;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = set_err(c, MFA_ERR_HIP, "sync failed");
  hipFree(pv);
  hipFree(pm);
  hipFree(pf);
  hipFree(pc);
  return rc ? rc : check_flags(c);
}

// Average durations (ms) of the three kernels of a rebuild (cell fill, cell
// sort, list build), each bracketed by HIP events on the context's stream,
// over `reps` rebuilds forced at the current state (which they leave as it
// is, the skin reference included).
static int time_rebuild_impl(mfa_ctx* c, int32_t reps, double* ms, int forced);

extern "C" int32_t mfa_time_rebuild(mfa_ctx* c, int32_t reps, double* ms) {
  return time_rebuild_impl(c, reps, ms, 1);
}

extern "C" int32_t mfa_time_rebuild_idle(mfa_ctx* c, int32_t reps, double* ms) {
  return time_rebuild_impl(c, reps, ms, 0);
}

static int time_rebuild_impl(mfa_ctx* c, int32_t reps, double* ms, int forced) {
  if (!c || !ms || reps < 1 || reps > 1024) return MFA_ERR_ARG;
  int rc = ensure_list(c);
  if (rc) return rc;
  hipEvent_t ev[4];
  for (int q = 0; q < 4; ++q) HIPCHK(c, hipEventCreate(&ev[q]));
  double tot[3] = {0.0, 0.0, 0.0};
  for (int it = 0; it < reps + 1 && rc == MFA_OK; ++it) {  // the first one warms up
    if (forced)
      hipLaunchKernelGGL(k_zero_counts, dim3(grid_of(c->n_cells)), dim3(MFA_BLOCK), 0, c->stream,
                         c->n_cells, c->cell_count);
    hipLaunchKernelGGL(k_set_flag, dim3(grid_of(c->n_rep)), dim3(MFA_BLOCK), 0, c->stream, c->flags,
                       c->n_rep, (int)F_REBUILD, forced);
    rc = launch_rebuild(c, ev);
    hipLaunchKernelGGL(k_set_flag, dim3(grid_of(c->n_rep)), dim3(MFA_BLOCK), 0, c->stream, c->flags,
                       c->n_rep, (int)F_REBUILD, 0);
    if (rc) break;
    if (hipEventSynchronize(ev[3]) != hipSuccess) {
      rc = set_err(c, MFA_ERR_HIP, "event synchronize failed");
      break;
    }
    for (int q = 0; q < 3 && it > 0; ++q) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, ev[q], ev[q + 1]) != hipSuccess) {
        rc = set_err(c, MFA_ERR_HIP, "event elapsed time failed");
        break;
      }
      tot[q] += t;
    }
  }
  for (int q = 0; q < 4; ++q) hipEventDestroy(ev[q]);
  if (rc) return rc;
  for (int q = 0; q < 3; ++q) ms[q] = tot[q] / reps;
  return check_flags(c);
}

// The k-space stage as the step runs it: ewald_k(1) and, with walls
// (SYMMETRY 0), the slab dipole correction.
static int launch_ewald_stage(mfa_ctx* c) { return launch_ewald(c, c->p.symmetry == 0 ? 1 : 0); }

extern "C" int32_t mfa_time_ewald(mfa_ctx* c, int32_t reps, double* ms_avg) {
  if (!c || !ms_avg || reps < 1 || reps > 4096) return MFA_ERR_ARG;
  if (!ewald_active(c)) return set_err(c, MFA_ERR_ARG, "no k-space Ewald stage (no charges)");
  // the launches add to F: keep F and put it back afterwards
  const size_t fb = 3 * (size_t)c->n * sizeof(double);
  double* keep = nullptr;
  if (hipMalloc(&keep, fb) != hipSuccess)
    return set_err(c, MFA_ERR_NOMEM, "hipMalloc of %zu bytes (F copy) failed", fb);
  int rc = MFA_OK;
  if (hipMemcpyAsync(keep, c->f, fb, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
    rc = set_err(c, MFA_ERR_HIP, "F copy failed");
  if (!rc) rc = time_launches(c, reps, launch_ewald_stage, ms_avg);
  if (hipMemcpyAsync(c->f, keep, fb, hipMemcpyDeviceToDevice, c->stream) != hipSuccess && !rc)
    rc = set_err(c, MFA_ERR_HIP, "F restore failed");
  if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = set_err(c, MFA_ERR_HIP, "sync failed");
  hipFree(keep);
  return rc ? rc : check_flags(c);
}

// Average durations (ms) of the five kernels of the k-space stage (tables,
// structure factor, its sum, forces, apply), each timed like mfa_time_force:
// `reps` back-to-back launches of that kernel alone (captured in one graph
// when graph steps are on) between two HIP events, after one full stage has
// filled its inputs; F is put back afterwards.
extern "C" int32_t mfa_time_ewald_parts(mfa_ctx* c, int32_t reps, double* ms) {
  if (!c || !ms || reps < 1 || reps > 4096) return MFA_ERR_ARG;
  if (!ewald_active(c)) return set_err(c, MFA_ERR_ARG, "no k-space Ewald stage (no charges)");
  const size_t fb = 3 * (size_t)c->n * sizeof(double);
  double* keep = nullptr;
  if (hipMalloc(&keep, fb) != hipSuccess)
    return set_err(c, MFA_ERR_NOMEM, "hipMalloc of %zu bytes (F copy) failed", fb);
  int rc = MFA_OK;
  if (hipMemcpyAsync(keep, c->f, fb, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
    rc = set_err(c, MFA_ERR_HIP, "F copy failed");
  if (!rc) rc = launch_ewald_stage(c);
  int (*parts[5])(mfa_ctx*) = {ew_tables, ew_rho, ew_sum, ew_force, ew_apply_stage};
  for (int q = 0; q < 5 && !rc; ++q) rc = time_launches(c, reps, parts[q], &ms[q]);
  if (hipMemcpyAsync(c->f, keep, fb, hipMemcpyDeviceToDevice, c->stream) != hipSuccess && !rc)
    rc = set_err(c, MFA_ERR_HIP, "F restore failed");
  if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = set_err(c, MFA_ERR_HIP, "sync failed");
  hipFree(keep);
  return rc ? rc : check_flags(c);
}

extern "C" int32_t mfa_copy_probe(int64_t bytes, double* ms) {
  if (bytes <= 0 || !ms) return set_err(nullptr, MFA_ERR_ARG, "bad copy size");
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = MFA_OK;
  if (hipMalloc(&a, (size_t)bytes) != hipSuccess || hipMalloc(&b, (size_t)bytes) != hipSuccess ||
      hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = set_err(nullptr, MFA_ERR_NOMEM, "copy probe allocation failed");
  } else if (hipMemsetAsync(a, 1, (size_t)bytes, st) != hipSuccess ||
             hipMemcpyAsync(b, a, (size_t)bytes, hipMemcpyDeviceToDevice, st) != hipSuccess ||
             hipEventRecord(e0, st) != hipSuccess ||
             hipMemcpyAsync(b, a, (size_t)bytes, hipMemcpyDeviceToDevice, st) != hipSuccess ||
             hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
    rc = set_err(nullptr, MFA_ERR_HIP, "copy probe failed");
  } else {
    float t = 0.f;
    hipEventElapsedTime(&t, e0, e1);
    *ms = t;
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st) hipStreamDestroy(st);
  if (a) hipFree(a);
  if (b) hipFree(b);
  return rc;
}

extern "C" int32_t mfa_stats(mfa_ctx* c, int64_t* out) {
  if (!c || !out) return MFA_ERR_ARG;
  HIPCHK(c, hipMemcpyAsync(c->h_flags, c->flags, F_NFLAGS * sizeof(int32_t), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  out[0] = c->h_flags[F_MAXNEIGH];
  out[1] = c->cap;
  out[2] = c->h_flags[F_ERROR];
  out[3] = c->h_flags[F_NREBUILD];
  return MFA_OK;
}

extern "C" int32_t mfa_timing(mfa_ctx* c, int32_t enable, double* force_ms_avg,
                              int64_t* force_launches, double* integrate_ms_avg,
                              int64_t* integrate_launches) {
  if (!c) return MFA_ERR_ARG;
  int rc = timed_flush(c, c->ev_force, c->ev_force_used, c->force_ms, c->force_n);
  if (rc) return rc;
  rc = timed_flush(c, c->ev_integ, c->ev_integ_used, c->integ_ms, c->integ_n);
  if (rc) return rc;
  if (force_ms_avg) *force_ms_avg = c->force_n ? c->force_ms / (double)c->force_n : 0.0;
  if (force_launches) *force_launches = c->force_n;
  if (integrate_ms_avg) *integrate_ms_avg = c->integ_n ? c->integ_ms / (double)c->integ_n : 0.0;
  if (integrate_launches) *integrate_launches = c->integ_n;
  c->timing = enable != 0;
  c->force_ms = 0.0;
  c->force_n = 0;
  c->integ_ms = 0.0;
  c->integ_n = 0;
  return MFA_OK;
}

// ---------------------------------------------------------------------------
// Adjoint of the force-and-integrate step (no counterpart in the reference).
//
// Step n of the window maps (r, v, a)_{n-1} to (r, v, a)_n:
//   r_n = r_{n-1} + dt v_{n-1} + dt^2/2 a_{n-1}     u_n = v_{n-1} + dt/2 a_{n-1}
//   a_n = F(r_n, u_n; theta, xi_n) / m              v_n = u_n + dt/2 a_n
// (pinned heads: v, a reset, so nothing flows through theirs).  Going
// backward, with adjoint vectors (lr, lv, la) of (r, v, a)_n:
//   lr += dJ_n/dr_n;  lu = lv;  lF = (la + dt/2 lv) / m;
//   lr += (dF/dr)^T lF;  lu += (dF/du)^T lF;  g_theta += (dF/dtheta)^T lF;
//   lv <- lu + dt lr;  la <- dt/2 lu + dt^2/2 lr.
// A central pair term F_ij = s(r^2, w) d (d = r_i - r_j, w = (u_i - u_j).d)
// contributes with L = (lF_i - lF_j).F_ij:
//   dL/dr_i = s Lam + (Lam.d) (2 s_r2 d + s_w du),   dL/du_i = (Lam.d) s_w d,
// and the negatives to j; the full list holds each pair in both directions,
// so each particle gathers its own terms and the parameter sums take half of
// every entry.  Both observables (brush height, top-wall pressure) are carried
// at once: every adjoint array is double[N][2][3].

#define ADJ_NO 2  // observables
#define ADJ_NP 5  // epsilon_LJ, sigma, k_FENE, R0, gamma_DPD
enum { AP_EPS = 0, AP_SIG, AP_K, AP_R0, AP_GAM };

struct AdjK {
  double inv_w;     // 1 / window length
  double inv_n2;    // 1 / number of type-2 beads
  double inv_area;  // 1 / (Lx Ly)
  double eps_in;    // common epsilon (epsil / 4)
  double sigma;     // common sigma
  double k_in, r0;  // FENE k and R0
  double gamma;     // friction(1)
};

// Top-wall force on a bead at height z (fluid_wall.f90:107-121) and its
// derivatives with respect to z and to sigma_wall.
__device__ __forceinline__ void wall_top_terms(double sw, double aw, double z, double zw,
                                               double& f, double& df_dz, double& df_dsw) {
  const double aabs = fabs(aw);
  const double iz = 1.0 / (zw - z);
  const double iz2 = iz * iz, iz4 = iz2 * iz2, iz5 = iz4 * iz, iz8 = iz4 * iz4;
  const double iz10 = iz8 * iz2, iz11 = iz10 * iz;
  const double sw2 = sw * sw, sw3 = sw2 * sw, sw8 = sw3 * sw3 * sw2, sw9 = sw8 * sw;
  f = -(9.0 * aabs) * sw9 * iz10 + (3.0 * aw) * sw3 * iz4;
  df_dz = -90.0 * aabs * sw9 * iz11 + 12.0 * aw * sw3 * iz5;
  df_dsw = -(81.0 * aabs) * sw8 * iz10 + (9.0 * aw) * sw2 * iz4;
}

// Observables of the forward pass at r_n: partial sums of z over type-2
// beads (slot 0) and of the top-wall force (slot 1).
__global__ __launch_bounds__(MFA_BLOCK) void k_adj_obs(const KParams* __restrict__ kb,
                                                       const int* __restrict__ rmap,
                                                       const double* __restrict__ r,
                                                       const int* __restrict__ type,
                                                       double* __restrict__ part) {
  MFA_REPLICA(blockIdx.x * blockDim.x);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double vals[MFA_NTERMS];
  for (int t = 0; t < MFA_NTERMS; ++t) vals[t] = 0.0;
  if (i - k.p0 < k.n) {
    const int ti = type[i] - 1;
    const double z = r[3 * (size_t)i + 2];
    if (ti == 1) vals[0] = z;
    double f, dz, ds;
    wall_top_terms(k.sigma_wall[ti], k.a_wall[ti], z, k.z_space_wall, f, dz, ds);
    vals[1] = f;
  }
  block_partials(vals, part);
}

// Seeds of step n and the force adjoint: lr += dJ_n/dr_n, lu = lv,
// lF = (la + dt/2 lv)/m (0 for pinned heads).  Slot AP_SIG of observable 1
// (MFA_NTERMS layout: o * ADJ_NP + p) gets the direct sigma term of the
// pressure through sigma_wall = (sigma_w + sigma)/2.
__global__ __launch_bounds__(MFA_BLOCK) void k_adj_pre(const KParams* __restrict__ kb,
                                                       const int* __restrict__ rmap, AdjK ak,
                                                       const double* __restrict__ r,
                                                       const int* __restrict__ type,
                                                       double* __restrict__ lr,
                                                       const double* __restrict__ lv,
                                                       const double* __restrict__ la,
                                                       double* __restrict__ lf,
                                                       double* __restrict__ part) {
  MFA_REPLICA(blockIdx.x * blockDim.x);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double vals[MFA_NTERMS];
  for (int t = 0; t < MFA_NTERMS; ++t) vals[t] = 0.0;
  if (i - k.p0 < k.n) {
    const int ti = type[i] - 1;
    const double z = r[3 * (size_t)i + 2];
    double f, dz, ds;
    wall_top_terms(k.sigma_wall[ti], k.a_wall[ti], z, k.z_space_wall, f, dz, ds);
    double* lri = lr + 6 * (size_t)i;
    if (ti == 1) lri[2] += ak.inv_w * ak.inv_n2;
    lri[5] += -ak.inv_w * ak.inv_area * dz;
    vals[ADJ_NP + AP_SIG] = -ak.inv_w * ak.inv_area * ds * 0.5;
    const bool head = is_head(k, i - k.p0);
    const double im = k.inv_mass_type[ti];
    for (int c = 0; c < 6; ++c) {
      const double lac = la[6 * (size_t)i + c] + k.half_dt * lv[6 * (size_t)i + c];
      lf[6 * (size_t)i + c] = head ? 0.0 : lac * im;
    }
  }
  block_partials(vals, part);
}

// lv <- lu + dt lr, la <- dt/2 lu + dt^2/2 lr (the adjoint of
// verlet_positions); nothing flows into a pinned head's v or a.
__global__ __launch_bounds__(MFA_BLOCK) void k_adj_post(const KParams* __restrict__ kb,
                                                        const int* __restrict__ rmap,
                                                        const double* __restrict__ lr,
                                                        double* __restrict__ lv,
                                                        double* __restrict__ la) {
  MFA_REPLICA(blockIdx.x * blockDim.x);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i - k.p0 >= k.n) return;
  const bool head = is_head(k, i - k.p0);
  for (int c = 0; c < 6; ++c) {
    const double lu = lv[6 * (size_t)i + c], lrc = lr[6 * (size_t)i + c];
    lv[6 * (size_t)i + c] = head ? 0.0 : lu + k.dt * lrc;
    la[6 * (size_t)i + c] = head ? 0.0 : k.half_dt * lu + k.half_dt_2 * lrc;
  }
}

// The force adjoint: lr_i += (dF/dr_i)^T lF, lu_i += (dF/du_i)^T lF over the
// pair, bond and wall terms of the force pass, FP64; parameter partials in
// slots o * ADJ_NP + p.  TPP lanes share a particle's list as in k_force64.
template <int TPP>
__global__ __launch_bounds__(MFA_BLOCK) void k_adj_force(
    const KParams* __restrict__ kb, const int* __restrict__ rmap, AdjK ak,
    const double* __restrict__ r, const double* __restrict__ u, const int* __restrict__ type,
    const int* __restrict__ perm, const unsigned* __restrict__ nl, const int* __restrict__ nn,
    int cap, const double* __restrict__ lf, double* __restrict__ lr, double* __restrict__ lu,
    const long long* __restrict__ d_step, double* __restrict__ part) {
  MFA_REPLICA(blockIdx.x * (MFA_BLOCK / TPP));
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int sp = gtid / TPP;  // sorted position
  const int sub = gtid % TPP;
  const int64_t step = *d_step;
  const bool active = sp - k.p0 < k.n;
  const int ss = active ? sp : k.p0;
  const int i = perm[ss];
  const int il = i - k.p0;
  const int ii = i;
  const double xi = r[3 * (size_t)ii + 0], yi = r[3 * (size_t)ii + 1], zi = r[3 * (size_t)ii + 2];
  const double uxi = u[3 * (size_t)ii + 0], uyi = u[3 * (size_t)ii + 1], uzi = u[3 * (size_t)ii + 2];
  const int ti = type[ii] - 1;
  double lfi[6];
  for (int c = 0; c < 6; ++c) lfi[c] = lf[6 * (size_t)ii + c];
  double gr[6], gu[6], gp[ADJ_NO * ADJ_NP];
  for (int c = 0; c < 6; ++c) gr[c] = gu[c] = 0.0;
  for (int c = 0; c < ADJ_NO * ADJ_NP; ++c) gp[c] = 0.0;

  if (active) {
    const int wn = nn[ss], cnt = wn & 0xffff, nup = wn >> 16;
    const unsigned* row = nl + (size_t)ss * cap;
    for (int q = sub; q < cnt; q += TPP) {
      const int j = perm[row[list_slot(q, nup, cnt, cap)] & MFA_JMASK];
      double d[3] = {xi - r[3 * (size_t)j + 0], yi - r[3 * (size_t)j + 1], zi - r[3 * (size_t)j + 2]};
      min_image(k, d[0], d[1], d[2]);
      const double r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
      const int tj = type[j] - 1;
      const bool in_lj = r2 < k.range_2[ti][tj];
      bool dpd = false;
      double rc = 0.0;
      if (k.dpd_on) {
        if (!k.dpd_cut_off_on) {
          dpd = in_lj;
          rc = sqrt(k.range_2[ti][tj]);
        } else {
          const bool excl = k.dpd_head_excl && (ti == 0 || tj == 0);
          dpd = !excl && r2 < k.r_cut_dpd_2;
          rc = sqrt(k.r_cut_dpd_2);
        }
      }
      if (!in_lj && !dpd) continue;
      double s = 0.0, s_r2 = 0.0, s_w = 0.0, s_eps = 0.0, s_sig = 0.0, s_gam = 0.0;
      const double du[3] = {uxi - u[3 * (size_t)j + 0], uyi - u[3 * (size_t)j + 1],
                            uzi - u[3 * (size_t)j + 2]};
      if (in_lj) {
        const double eps = k.epsil[ti][tj];
        const double inv_r2 = 1.0 / r2;
        const double r61 = k.sigma_2[ti][tj] * inv_r2;
        const double r6 = (r61 * r61) * r61, r12 = r6 * r6;
        s += eps * (12.0 * r12 - 6.0 * r6) * inv_r2;
        s_r2 += eps * (-84.0 * r12 + 24.0 * r6) * inv_r2 * inv_r2;
        if (eps != 0.0) s_eps += 4.0 * (12.0 * r12 - 6.0 * r6) * inv_r2;
        s_sig += eps * (144.0 * r12 - 36.0 * r6) * inv_r2 / ak.sigma;
      }
      if (dpd) {
        const double rr = sqrt(r2);
        const double wr = 1.0 - rr / rc;
        const double xi_n = pair_normal_f64(k.seed, step, il, j - k.p0);
        const double w = (du[0] * d[0] + du[1] * d[1]) + du[2] * d[2];
        const double A = k.sig * wr * xi_n / rr;          // random term / r
        const double B = -k.friction1 * wr * wr * w / r2;  // dissipative term / r
        s += A + B;
        s_r2 += -k.sig * xi_n / (2.0 * rr * r2) + k.friction1 * w * wr / (r2 * r2);
        s_w += -k.friction1 * wr * wr / r2;
        s_gam += A / (2.0 * k.friction1) + B / k.friction1;
        if (!k.dpd_cut_off_on) {  // the DPD range is the LJ range: it moves with sigma
          const double dwr = (rr / rc) / ak.sigma;
          s_sig += k.sig * xi_n * dwr / rr - 2.0 * k.friction1 * wr * dwr * w / r2;
        }
      }
      for (int o = 0; o < ADJ_NO; ++o) {
        double lam[3];
        for (int c = 0; c < 3; ++c) lam[c] = lfi[3 * o + c] - lf[6 * (size_t)j + 3 * o + c];
        const double ld = (lam[0] * d[0] + lam[1] * d[1]) + lam[2] * d[2];
        for (int c = 0; c < 3; ++c) {
          gr[3 * o + c] += s * lam[c] + ld * (2.0 * s_r2 * d[c] + s_w * du[c]);
          gu[3 * o + c] += ld * s_w * d[c];
        }
        gp[o * ADJ_NP + AP_EPS] += 0.5 * ld * s_eps;
        gp[o * ADJ_NP + AP_SIG] += 0.5 * ld * s_sig;
        gp[o * ADJ_NP + AP_GAM] += 0.5 * ld * s_gam;
      }
    }
  }

  if (active && sub == 0) {
    // FENE bonds (intra_molec.f90:16-139): F_hi = rd d, F_lo = -rd d with
    // d = r_lo - r_hi, rd = k / (1 - x), x = r^2/R0^2 clamped at 0.98.
    int base, len;
    chain_of(k, il, base, len);
    for (int side = 0; side < 2; ++side) {
      const int lol = side == 0 ? il - 1 : il;
      if (len < 2 || lol < base || lol + 1 >= base + len) continue;
      const int lo = k.p0 + lol, hi = lo + 1;
      double d[3] = {r[3 * (size_t)lo + 0] - r[3 * (size_t)hi + 0],
                     r[3 * (size_t)lo + 1] - r[3 * (size_t)hi + 1],
                     r[3 * (size_t)lo + 2] - r[3 * (size_t)hi + 2]};
      min_image(k, d[0], d[1], d[2]);
      const double r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
      double x = r2 * k.inv_r_chain_2;
      const double clamp = (double)0.98f;
      const bool free = x < clamp;
      x = free ? x : clamp;
      const double rd = (k.k_chain / (1.0 - x)) * k.inv_r_chain_2;
      const double rd_r2 = free ? rd / (1.0 - x) * k.inv_r_chain_2 : 0.0;
      const double sg = side == 0 ? -1.0 : 1.0;  // i is hi (side 0) or lo (side 1)
      for (int o = 0; o < ADJ_NO; ++o) {
        double lam[3];
        for (int c = 0; c < 3; ++c)
          lam[c] = lf[6 * (size_t)hi + 3 * o + c] - lf[6 * (size_t)lo + 3 * o + c];
        const double ld = (lam[0] * d[0] + lam[1] * d[1]) + lam[2] * d[2];
        for (int c = 0; c < 3; ++c) gr[3 * o + c] += sg * (rd * lam[c] + 2.0 * ld * rd_r2 * d[c]);
        if (side == 1) {  // each bond's parameter terms once, from its lo end
          gp[o * ADJ_NP + AP_K] += ld * rd / ak.k_in;
          if (free) gp[o * ADJ_NP + AP_R0] += ld * (-2.0 * x * rd / ((1.0 - x) * ak.r0));
        }
      }
    }
    if (k.symmetry == 0) {
      // 9-3 walls (fluid_wall.f90:81-121): F_z(z) of both walls
      const double sw = k.sigma_wall[ti], aw = k.a_wall[ti], aabs = fabs(aw);
      const double iz = 1.0 / zi, iz2 = iz * iz, iz4 = iz2 * iz2, iz5 = iz4 * iz;
      const double iz8 = iz4 * iz4, iz10 = iz8 * iz2, iz11 = iz10 * iz;
      const double sw2 = sw * sw, sw3 = sw2 * sw, sw8 = sw3 * sw3 * sw2, sw9 = sw8 * sw;
      const double db_dz = -90.0 * aabs * sw9 * iz11 + 12.0 * aw * sw3 * iz5;
      const double db_dsw = (81.0 * aabs) * sw8 * iz10 - (9.0 * aw) * sw2 * iz4;
      double ft, dt_dz, dt_dsw;
      wall_top_terms(sw, aw, zi, k.z_space_wall, ft, dt_dz, dt_dsw);
      for (int o = 0; o < ADJ_NO; ++o) {
        gr[3 * o + 2] += lfi[3 * o + 2] * (db_dz + dt_dz);
        gp[o * ADJ_NP + AP_SIG] += lfi[3 * o + 2] * (db_dsw + dt_dsw) * 0.5;
      }
    }
  }

  for (int c = 0; c < 6; ++c) {
    gr[c] = group_sum<TPP>(gr[c]);
    gu[c] = group_sum<TPP>(gu[c]);
  }
  if (active && sub == 0) {
    for (int c = 0; c < 6; ++c) {
      lr[6 * (size_t)i + c] += gr[c];
      lu[6 * (size_t)i + c] += gu[c];
    }
  }
  double vals[MFA_NTERMS];
  for (int t = 0; t < MFA_NTERMS; ++t) vals[t] = t < ADJ_NO * ADJ_NP ? gp[t] : 0.0;
  block_partials(vals, part);
}

// Device buffers of one adjoint window.
struct AdjBufs {
  double *ck_r = nullptr, *ck_v = nullptr, *ck_a = nullptr;  // checkpoints [nseg][N][3]
  int* ck_mic = nullptr;
  // the neighbour list in effect at each checkpoint (rows, counts, skin
  // reference, cells), so a recomputed segment repeats the forward bit for bit
  unsigned* ck_nl = nullptr;
  int *ck_nn = nullptr, *ck_cell = nullptr, *ck_perm = nullptr, *ck_sidx = nullptr;
  double* ck_rref = nullptr;
  double *sg_r = nullptr, *sg_u = nullptr;  // recomputed segment [K][N][3]
  int* sg_mic = nullptr;
  double *end_r = nullptr, *end_v = nullptr, *end_a = nullptr, *end_f = nullptr;
  int* end_mic = nullptr;
  double *lr = nullptr, *lv = nullptr, *la = nullptr, *lf = nullptr;  // [N][2][3]
  double* obs = nullptr;   // [W][MFA_NTERMS]
  double* grad = nullptr;  // [W][2][MFA_NTERMS]
  ~AdjBufs() {
    void* ptrs[] = {ck_r, ck_v, ck_a, ck_mic, ck_nl, ck_nn, ck_cell, ck_perm, ck_sidx, ck_rref, sg_r, sg_u, sg_mic,
                    end_r, end_v, end_a, end_f, end_mic, lr, lv, la, lf, obs, grad};
    for (void* q : ptrs)
      if (q) hipFree(q);
  }
};

template <int TPP>
static void launch_adj_force_t(mfa_ctx* c, const AdjK& ak, const AdjBufs& b) {
  hipLaunchKernelGGL((k_adj_force<TPP>), dim3(grid_of(c->n * TPP)), dim3(MFA_BLOCK), 0, c->stream,
                     c->d_kp, c->d_rmap, ak, c->r, c->v, c->type, c->perm, c->nl, c->nn, c->cap,
                     b.lf, b.lr, b.lv, c->d_step, c->part);
}

static int launch_adj_force(mfa_ctx* c, const AdjK& ak, const AdjBufs& b) {
  switch (c->tpp) {
    case 1: launch_adj_force_t<1>(c, ak, b); break;
    case 2: launch_adj_force_t<2>(c, ak, b); break;
    case 4: launch_adj_force_t<4>(c, ak, b); break;
    case 8: launch_adj_force_t<8>(c, ak, b); break;
    case 16: launch_adj_force_t<16>(c, ak, b); break;
    default: return set_err(c, MFA_ERR_ARG, "threads per particle %d unsupported", c->tpp);
  }
  HIPCHK(c, hipGetLastError());
  return MFA_OK;
}

static int d2d(mfa_ctx* c, void* dst, const void* src, size_t bytes) {
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  return MFA_OK;
}

// One forward step of the window (verlet_positions + skin test, rebuild when
// due, force pass, verlet_velocities) at step key `step`; with `obs_row` the
// observables at r_n go to that row, with seg_* the force-pass inputs
// (r_n, u_n, mic_n) are saved.
static int adj_forward_step(mfa_ctx* c, int64_t step, double* obs_row, double* seg_r, double* seg_u,
                            int* seg_mic) {
  const size_t n3 = (size_t)c->n * 3;
  int rc = set_step(c, step);
  if (rc) return rc;
  const int g = grid_of(c->n);
  hipLaunchKernelGGL(k_verlet_positions, dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->d_rmap,
                     c->r, c->v, c->a, c->mic, c->r_ref, c->flags, 1, c->n_cells, c->cell_count,
                     c->type, c->cell_of, c->sidx, c->pos4, c->vel4);
  HIPCHK(c, hipGetLastError());
  rc = launch_rebuild(c);
  if (rc) return rc;
  if (seg_r) {
    if ((rc = d2d(c, seg_r, c->r, n3 * sizeof(double))) || (rc = d2d(c, seg_u, c->v, n3 * sizeof(double))) ||
        (rc = d2d(c, seg_mic, c->mic, n3 * sizeof(int))))
      return rc;
  }
  rc = launch_force(c, TM_ALL);
  if (rc) return rc;
  if (obs_row) {
    hipLaunchKernelGGL(k_adj_obs, dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->d_rmap, c->r,
                       c->type, c->part);
    hipLaunchKernelGGL(k_reduce_partials, dim3(MFA_NTERMS), dim3(MFA_BLOCK), 0, c->stream, c->part, 0,
                       g, obs_row);
    HIPCHK(c, hipGetLastError());
  }
  hipLaunchKernelGGL(k_verlet_velocities, dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->d_rmap,
                     c->f, c->v, c->a, c->type, c->part, 0, c->sidx, c->vel4, nullptr);
  HIPCHK(c, hipGetLastError());
  return MFA_OK;
}

extern "C" int32_t mfa_adjoint_window(mfa_ctx* c, int64_t first_step, int32_t n_steps,
                                      int32_t checkpoint_every, int32_t want_grad,
                                      mfa_adjoint_out* out) {
  if (!c || !out || n_steps < 1 || checkpoint_every < 1)
    return set_err(c, MFA_ERR_ARG, "adjoint: bad arguments");
  const mfa_params& p = c->p;
  if (c->n_rep != 1) return set_err(c, MFA_ERR_ARG, "adjoint: one replica per context");
  if (p.coulomb || p.bending) return set_err(c, MFA_ERR_ARG, "adjoint: charges and bending are not covered");
  if (p.fso_time > 0.0) return set_err(c, MFA_ERR_ARG, "adjoint: the switch-on ramp is not covered");
  if (c->dpd_vv) return set_err(c, MFA_ERR_ARG, "adjoint: DPD_VV is not covered");
  if (p.thermostat == 1) return set_err(c, MFA_ERR_ARG, "adjoint: the Langevin thermostat is not covered");
  if (p.symmetry != 0) return set_err(c, MFA_ERR_ARG, "adjoint: the observables need the walls (SYMMETRY 0)");
  // theta: the common epsilon and sigma of the tables
  AdjK ak;
  memset(&ak, 0, sizeof ak);
  ak.eps_in = 0.0;
  ak.sigma = 0.0;
  for (int t = 0; t < MFA_NTYPE; ++t)
    for (int q = 0; q < MFA_NTYPE; ++q) {
      const double e = p.epsil[t][q] / 4.0, sg = p.sigma[t][q];
      if (e != 0.0) {
        if (ak.eps_in != 0.0 && e != ak.eps_in) return set_err(c, MFA_ERR_ARG, "adjoint: epsilon differs between types");
        ak.eps_in = e;
      }
      if (sg != 0.0) {
        if (ak.sigma != 0.0 && sg != ak.sigma) return set_err(c, MFA_ERR_ARG, "adjoint: sigma differs between types");
        ak.sigma = sg;
      }
    }
  if (ak.sigma == 0.0) return set_err(c, MFA_ERR_ARG, "adjoint: no sigma");
  ak.r0 = sqrt(p.r_chain_2);
  ak.k_in = p.k_chain * p.inv_r_chain_2;
  ak.gamma = p.friction1;
  const int W = n_steps, K = checkpoint_every < n_steps ? checkpoint_every : n_steps;
  const int nseg = (W + K - 1) / K;
  const size_t n = (size_t)c->n, n3 = n * 3;
  AdjBufs b;
#define AALLOC(ptr, bytes)                                                                 \
  do {                                                                                     \
    if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess)                                  \
      return set_err(c, MFA_ERR_NOMEM, "adjoint: hipMalloc of %zu bytes failed", (size_t)(bytes)); \
  } while (0)
  AALLOC(b.obs, (size_t)W * MFA_NTERMS * sizeof(double));
  AALLOC(b.ck_r, (size_t)nseg * n3 * sizeof(double));
  AALLOC(b.ck_v, (size_t)nseg * n3 * sizeof(double));
  AALLOC(b.ck_a, (size_t)nseg * n3 * sizeof(double));
  AALLOC(b.ck_mic, (size_t)nseg * n3 * sizeof(int));
  if (want_grad) {
    AALLOC(b.ck_nl, (size_t)nseg * n * c->cap * sizeof(unsigned));
    AALLOC(b.ck_nn, (size_t)nseg * n * sizeof(int));
    AALLOC(b.ck_cell, (size_t)nseg * n * sizeof(int));
    AALLOC(b.ck_perm, (size_t)nseg * n * sizeof(int));
    AALLOC(b.ck_sidx, (size_t)nseg * n * sizeof(int));
    AALLOC(b.ck_rref, (size_t)nseg * n3 * sizeof(double));
    AALLOC(b.sg_r, (size_t)K * n3 * sizeof(double));
    AALLOC(b.sg_u, (size_t)K * n3 * sizeof(double));
    AALLOC(b.sg_mic, (size_t)K * n3 * sizeof(int));
    AALLOC(b.end_r, n3 * sizeof(double));
    AALLOC(b.end_v, n3 * sizeof(double));
    AALLOC(b.end_a, n3 * sizeof(double));
    AALLOC(b.end_f, n3 * sizeof(double));
    AALLOC(b.end_mic, n3 * sizeof(int));
    AALLOC(b.lr, n * 6 * sizeof(double));
    AALLOC(b.lv, n * 6 * sizeof(double));
    AALLOC(b.la, n * 6 * sizeof(double));
    AALLOC(b.lf, n * 6 * sizeof(double));
    AALLOC(b.grad, (size_t)W * 2 * MFA_NTERMS * sizeof(double));
  }
#undef AALLOC
  int rc = ensure_list(c);
  if (rc) return rc;
  hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(MFA_BLOCK), 0, c->stream, c->flags, 1,
                     (int)F_NREBUILD, 0);
  // forward pass with checkpoints
  for (int s = 1; s <= W; ++s) {
    if ((s - 1) % K == 0) {
      const size_t off = (size_t)((s - 1) / K) * n3;
      if ((rc = d2d(c, b.ck_r + off, c->r, n3 * sizeof(double))) ||
          (rc = d2d(c, b.ck_v + off, c->v, n3 * sizeof(double))) ||
          (rc = d2d(c, b.ck_a + off, c->a, n3 * sizeof(double))) ||
          (rc = d2d(c, b.ck_mic + off, c->mic, n3 * sizeof(int))))
        return rc;
      if (want_grad) {
        const size_t sg = (size_t)((s - 1) / K);
        if ((rc = d2d(c, b.ck_nl + sg * n * c->cap, c->nl, n * c->cap * sizeof(unsigned))) ||
            (rc = d2d(c, b.ck_nn + sg * n, c->nn, n * sizeof(int))) ||
            (rc = d2d(c, b.ck_cell + sg * n, c->cell_of, n * sizeof(int))) ||
            (rc = d2d(c, b.ck_perm + sg * n, c->perm, n * sizeof(int))) ||
            (rc = d2d(c, b.ck_sidx + sg * n, c->sidx, n * sizeof(int))) ||
            (rc = d2d(c, b.ck_rref + off, c->r_ref, n3 * sizeof(double))))
          return rc;
      }
    }
    rc = adj_forward_step(c, first_step + s - 1, b.obs + (size_t)(s - 1) * MFA_NTERMS, nullptr,
                          nullptr, nullptr);
    if (rc) return rc;
  }
  rc = check_flags(c);
  if (rc) return rc;
  std::vector<double> h_obs((size_t)W * MFA_NTERMS);
  HIPCHK(c, hipMemcpyAsync(h_obs.data(), b.obs, h_obs.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  memset(out, 0, sizeof *out);
  out->n_rebuilds = c->h_flags[F_NREBUILD];
  int n2 = 0;
  {
    std::vector<int32_t> ht(n);
    HIPCHK(c, hipMemcpy(ht.data(), c->type, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int i = 0; i < p.n_part; ++i) n2 += ht[(size_t)i] == 2;
  }
  if (n2 == 0) return set_err(c, MFA_ERR_ARG, "adjoint: no type-2 beads for the brush height");
  ak.inv_w = 1.0 / (double)W;
  ak.inv_n2 = 1.0 / (double)n2;
  ak.inv_area = 1.0 / (p.boundary[0] * p.boundary[1]);
  {
    double zs = 0.0, fs = 0.0;
    for (int s = 0; s < W; ++s) {
      zs += h_obs[(size_t)s * MFA_NTERMS + 0] * ak.inv_n2;
      fs += -h_obs[(size_t)s * MFA_NTERMS + 1] * ak.inv_area;
    }
    out->height = zs * ak.inv_w;
    out->p_wall = fs * ak.inv_w;
  }
  if (!want_grad) return MFA_OK;

  // keep the end state, then run the segments backward
  if ((rc = d2d(c, b.end_r, c->r, n3 * sizeof(double))) || (rc = d2d(c, b.end_v, c->v, n3 * sizeof(double))) ||
      (rc = d2d(c, b.end_a, c->a, n3 * sizeof(double))) || (rc = d2d(c, b.end_f, c->f, n3 * sizeof(double))) ||
      (rc = d2d(c, b.end_mic, c->mic, n3 * sizeof(int))))
    return rc;
  HIPCHK(c, hipMemsetAsync(b.lr, 0, n * 6 * sizeof(double), c->stream));
  HIPCHK(c, hipMemsetAsync(b.lv, 0, n * 6 * sizeof(double), c->stream));
  HIPCHK(c, hipMemsetAsync(b.la, 0, n * 6 * sizeof(double), c->stream));
  HIPCHK(c, hipMemsetAsync(b.lf, 0, n * 6 * sizeof(double), c->stream));
  const int g = grid_of(c->n);
  const int gf = grid_of(c->n * c->tpp);
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int s0 = seg * K, s1 = s0 + K < W ? s0 + K : W;
    const size_t off = (size_t)seg * n3;
    if ((rc = d2d(c, c->r, b.ck_r + off, n3 * sizeof(double))) ||
        (rc = d2d(c, c->v, b.ck_v + off, n3 * sizeof(double))) ||
        (rc = d2d(c, c->a, b.ck_a + off, n3 * sizeof(double))) ||
        (rc = d2d(c, c->mic, b.ck_mic + off, n3 * sizeof(int))) ||
        (rc = d2d(c, c->nl, b.ck_nl + (size_t)seg * n * c->cap, n * c->cap * sizeof(unsigned))) ||
        (rc = d2d(c, c->nn, b.ck_nn + (size_t)seg * n, n * sizeof(int))) ||
        (rc = d2d(c, c->cell_of, b.ck_cell + (size_t)seg * n, n * sizeof(int))) ||
        (rc = d2d(c, c->perm, b.ck_perm + (size_t)seg * n, n * sizeof(int))) ||
        (rc = d2d(c, c->sidx, b.ck_sidx + (size_t)seg * n, n * sizeof(int))) ||
        (rc = d2d(c, c->r_ref, b.ck_rref + off, n3 * sizeof(double))))
      return rc;
    // a restored list: the half-list pass must find its mirror positions again
    hipLaunchKernelGGL(k_set_flag, dim3(grid_of(c->n_rep)), dim3(MFA_BLOCK), 0, c->stream, c->flags,
                       c->n_rep, (int)F_MIRROR, 1);
    for (int s = s0 + 1; s <= s1; ++s) {
      const size_t so = (size_t)(s - s0 - 1) * n3;
      rc = adj_forward_step(c, first_step + s - 1, nullptr, b.sg_r + so, b.sg_u + so, b.sg_mic + so);
      if (rc) return rc;
    }
    for (int s = s1; s > s0; --s) {
      const size_t so = (size_t)(s - s0 - 1) * n3;
      if ((rc = d2d(c, c->r, b.sg_r + so, n3 * sizeof(double))) ||
          (rc = d2d(c, c->v, b.sg_u + so, n3 * sizeof(double))) ||
          (rc = d2d(c, c->mic, b.sg_mic + so, n3 * sizeof(int))))
        return rc;
      // the list must hold the pairs at r_n: rebuild when r_n left the skin
      hipLaunchKernelGGL(k_zero_counts, dim3(grid_of(c->n_cells)), dim3(MFA_BLOCK), 0, c->stream,
                         c->n_cells, c->cell_count);
      hipLaunchKernelGGL(k_check_skin, dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->d_rmap,
                         c->r, c->r_ref, c->mic, c->flags);
      rc = launch_rebuild(c);
      if (rc) return rc;
      hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(MFA_BLOCK), 0, c->stream, c->flags, 1,
                         (int)F_REBUILD, 0);
      rc = set_step(c, first_step + s - 1);
      if (rc) return rc;
      double* grow = b.grad + (size_t)(s - 1) * 2 * MFA_NTERMS;
      hipLaunchKernelGGL(k_adj_pre, dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->d_rmap, ak,
                         c->r, c->type, b.lr, b.lv, b.la, b.lf, c->part);
      hipLaunchKernelGGL(k_reduce_partials, dim3(MFA_NTERMS), dim3(MFA_BLOCK), 0, c->stream, c->part,
                         0, g, grow);
      rc = launch_adj_force(c, ak, b);
      if (rc) return rc;
      hipLaunchKernelGGL(k_reduce_partials, dim3(MFA_NTERMS), dim3(MFA_BLOCK), 0, c->stream, c->part,
                         0, gf, grow + MFA_NTERMS);
      hipLaunchKernelGGL(k_adj_post, dim3(g), dim3(MFA_BLOCK), 0, c->stream, c->d_kp, c->d_rmap, b.lr,
                         b.lv, b.la);
      HIPCHK(c, hipGetLastError());
    }
  }
  rc = check_flags(c);
  if (rc) return rc;
  std::vector<double> h_grad((size_t)W * 2 * MFA_NTERMS);
  HIPCHK(c, hipMemcpyAsync(h_grad.data(), b.grad, h_grad.size() * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  // restore the end state of the window
  if ((rc = d2d(c, c->r, b.end_r, n3 * sizeof(double))) || (rc = d2d(c, c->v, b.end_v, n3 * sizeof(double))) ||
      (rc = d2d(c, c->a, b.end_a, n3 * sizeof(double))) || (rc = d2d(c, c->f, b.end_f, n3 * sizeof(double))) ||
      (rc = d2d(c, c->mic, b.end_mic, n3 * sizeof(int))))
    return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->need_rebuild = true;
  for (int s = 0; s < W; ++s)
    for (int row = 0; row < 2; ++row)
      for (int q = 0; q < ADJ_NP; ++q) {
        const double* gr = &h_grad[((size_t)s * 2 + row) * MFA_NTERMS];
        out->d_height[q] += gr[q];
        out->d_p_wall[q] += gr[ADJ_NP + q];
      }
  return MFA_OK;
}
