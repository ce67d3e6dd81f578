"""Named workloads as raw inputs (the numbers system_input / mfa_input and the
control_simulation.h switches would hold), plus small parity cases.

Workloads (sizes as stated in the benchmark definitions):
  brush9k   DPD polymer brush, 10x10x20 box, 100 chains x 30 + 6000 solvent
  melt100k  Kremer-Grest melt, 2000 chains x 50 at density 0.85, periodic
The choices the workload text leaves open are fixed here:
  - heads at z=1.0 with x,y uniform (reference style, gen_brush.f90:9-30);
  - 9-3 walls A_w=-2, sigma_w=1 (repulsive, as the shipped system_input:11);
  - solvent is an ideal DPD fluid: epsilon(solvent, .) = 0;
  - list skin 0.3 (the reference hard-codes 0.4, init_params.f90:800);
  - the melt's DPD range equals its LJ cut-off 2.5 (no DPD_CUT_OFF);
  - generated beads keep 0.8 sigma apart (melt) or 0.7 sigma (the brush at
    grafting density 1 cannot be packed at 0.8).
"""
import numpy as np

from . import lib

NT = lib.NTYPE


def base_inputs():
    """mfa_input / system_input defaults: unit masses, sigma and epsilon."""
    x = lib.Inputs()
    x.s_time = 2  # > 1: no force switch-on ramp (init_config.f90:264-281)
    x.n_relax = 0
    x.n_obser = 0
    x.n_safe = 1000
    x.temp = 1.0
    x.x_space = 1.0
    x.y_space = 1.0
    for t in range(NT):
        x.mass_type[t] = 1.0
        x.sigma_ii[t] = 1.0
        x.friction[t] = 1.0
        for u in range(NT):
            x.epsil[t][u] = 1.0
    x.k_chain = 30.0
    x.r_chain = 1.5
    x.a_w = -2.0
    x.sigma_w = 1.0
    x.a_w4 = -2.0
    x.sigma_w4 = 1.0
    x.inter4 = 1
    x.system = 2
    x.symmetry = 1
    x.solvent = 1
    x.thermostat = 0
    x.iseed = 1
    return x


def _ideal_solvent(x, solvent_type=3):
    s = solvent_type - 1
    for t in range(NT):
        x.epsil[s][t] = 0.0
        x.epsil[t][s] = 0.0


def brush(n_chain, n_mon, n_solvent, lx, ly, lz, x_space=1.0, dt=0.005, gamma=4.5,
          dpd_cut_off=1.0, skin=0.3, seed=1001, thermostat=0):
    """Brush grafted on the z=0 wall in a channel with 9-3 walls (SYSTEM 3,
    SYMMETRY 0, WALL 2), ideal DPD solvent as single-bead 'melt' chains."""
    x = base_inputs()
    x.system = 3
    x.symmetry = 0
    x.n_mon, x.n_chain = n_mon, n_chain
    x.n_mon_d, x.n_chain_d = 1, n_solvent
    x.x_space = x_space
    x.y_space = x_space
    x.n_cell_w_x = int(round(lx / x_space))
    x.n_cell_w_y = int(round(ly / x_space))
    x.z_space_wall = lz
    x.dt = dt
    x.friction[0] = gamma
    x.dpd_cut_off = dpd_cut_off
    x.skin = skin
    x.iseed = seed
    x.thermostat = thermostat
    _ideal_solvent(x)
    return x


def melt(n_chain, n_mon, density, lj_rc=2.5, dt=0.01, gamma=1.0, skin=0.3, seed=2002,
         thermostat=0, n_cell=None, switch_on=False):
    """Periodic Kremer-Grest melt (SYSTEM 2, SYMMETRY 1), chains as melt
    ('droplet') chains of type 3, LJ cut-off lj_rc (survey patch P3)."""
    x = base_inputs()
    x.system = 2
    x.symmetry = 1
    n = n_chain * n_mon
    L = (n / density) ** (1.0 / 3.0)
    nc = n_cell if n_cell else max(8, int(L))
    x.n_cell_w_x = nc
    x.n_cell_w_y = nc
    x.x_space = L / nc
    x.y_space = L / nc
    x.z_space_wall = x.x_space * nc
    x.n_mon, x.n_chain = 0, 0
    x.n_mon_d, x.n_chain_d = n_mon, n_chain
    x.dt = dt
    x.friction[0] = gamma
    x.lj_rc = lj_rc
    x.skin = skin
    x.iseed = seed
    x.thermostat = thermostat
    if switch_on:
        # fresh start with FORCE_SWITCH_ON: the reference ramps the minimum
        # pair distance over 40 tau (init_force_switch_on.f90:15-26)
        x.s_time = 1
        x.force_switch_on = 1
    return x


def pe_brush(n_chain, n_mon, n_solvent, lx, ly, lz, x_space=1.0, dt=0.005, gamma=4.5,
             alpha_c=0.81, n_kmax=10, rc_coul=3.0, compat=False, skin=0.3, seed=3003):
    """Polyelectrolyte brush with Ewald (SYSTEM 3, PARTICLE_4): brush monomers
    type 2, every third bead type 4 with charge -1, counterions type 3 (+1)
    and neutral solvent type 2 as single beads; WCA everywhere, DPD cut-off
    1.0.  alpha_c is the reference's alpha_c (erfc(sqrt(alpha_c) r)); with
    `compat` the real space runs inside the LJ range with the reference's
    1/pi factor (ewald_real.f90, init_config.f90:230)."""
    x = brush(n_chain, n_mon, 0, lx, ly, lz, x_space=x_space, dt=dt, gamma=gamma, skin=skin,
              seed=seed)
    n_ci = n_chain * sum(1 for m in range(1, n_mon) if (m + 1) % 3 == 0)
    x.n_mon_d, x.n_chain_d = 1, n_ci + n_solvent
    x.particle_4 = 1
    x.inter4 = 1
    for t in range(NT):
        for u in range(NT):
            x.epsil[t][u] = 1.0
    x.charge[0], x.charge[1], x.charge[2], x.charge[3] = 0.0, 0.0, 1.0, -1.0
    x.alpha_c = alpha_c
    x.n_kmax = n_kmax
    x.ewald_compat = 1 if compat else 0
    x.rc_coul = 0.0 if compat else rc_coul
    return x


def stiff(x, k_bend, alpha_eq):
    """Adds chain stiffness (BENDING): V = k_bend/2 (alpha - alpha_eq)^2 per
    angle of the brush chains (bending.f90:33-127)."""
    x.bending = 1
    x.k_bend = k_bend
    x.alpha_eq = alpha_eq
    return x


def ensemble_replica(q, base_seed=4004, lz=12.5, solvent_density=3.0):
    """Replica q of the uneven brush ensemble: grafting density uniform in
    [0.05, 0.5], chain length uniform in {20..100}, lateral edge uniform in
    {8..40} sigma, channel height lz, total density 3 with the solvent
    filling what the brush leaves (so sizes run from ~2.4k to ~60k beads);
    chains are capped at a brush density of 0.5 so they fit the channel.
    Seed base_seed*1000 + q.  Fresh start with the force switch-on ramp
    (the generated beads keep 0.7 sigma apart)."""
    rng = np.random.default_rng(base_seed * 1000 + q)
    rho_g = rng.uniform(0.05, 0.5)
    n_mon = int(rng.integers(20, 101))
    L = int(rng.integers(8, 41))
    area = float(L * L)
    n_chain = max(1, min(int(round(rho_g * area)), int(0.5 * area * lz / n_mon)))
    n_chain = max(2, n_chain - n_chain % 2)  # even (init_params.f90:881-887)
    n_total = int(round(solvent_density * area * lz))
    n_solv = max(0, n_total - n_chain * n_mon)
    x = brush(n_chain, n_mon, n_solv, float(L), float(L), lz, seed=base_seed * 1000 + q)
    x.s_time = 1
    x.force_switch_on = 1
    return x


# List skins the autotuner tries where a workload's spec leaves the skin open
# (those wider than the cells allow are skipped).
SKIN_CANDIDATES = (0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9)
# candidate cell widths of the cell-size autotuner, as multiples of the
# reference rule's minimum width r_cut_max + 0.4 (make_binning_boxes.f90:42-47):
# the reference's grid first, then coarser ones (which hold wider skins)
CELL_WIDTH_FACTORS = (1.0, 1.08, 1.16, 1.25)


def cell_width_candidates(p):
    rule = p.r_cut_max + float(np.float32(0.4))
    return [rule * f for f in CELL_WIDTH_FACTORS]

WORKLOADS = {
    # plain FP64 DPD polymer brush (the CPU-runnable case)
    "brush9k": dict(kind=1, steps=1000, min_dist=0.7, skin_fixed=True,  # its spec: skin 0.3
                    make=lambda: brush(100, 30, 6000, 10.0, 10.0, 20.0, seed=1001)),
    # Kremer-Grest melt, the headline workload
    "melt100k": dict(kind=2, steps=5000, min_dist=0.8,
                     make=lambda: melt(2000, 50, 0.85, seed=2002)),
    # polyelectrolyte brush with Ewald: 400 chains x 40 (charge -1 on every
    # third bead), 5200 counterions, 5800 neutral solvent = 27,000 beads in
    # 63 x 63 x 40; alpha_c = 0.81, kmax = 10 (3997 k vectors), rc_coul = 3
    # adjoint window: DPD brush of 500 chains x 40 at grafting density 0.2
    # plus 30,000 solvent = 50,000 beads in 50 x 50 x 30
    "adjoint50k": dict(kind=1, steps=200, min_dist=0.7,
                       make=lambda: brush(500, 40, 30000, 50.0, 50.0, 30.0, seed=5005)),
    "pe_brush27k": dict(kind=3, steps=2000, min_dist=0.7,
                        make=lambda: pe_brush(400, 40, 5800, 63.0, 63.0, 40.0, n_kmax=10,
                                              alpha_c=0.81, rc_coul=3.0, seed=3003)),
}


def prepare_brush(inputs, kind, min_dist=0.7, precision=lib.MFA_MIXED):
    """Seeded brush start relaxed by the reference's start-up procedure: the
    FORCE_SWITCH_ON ramp of a fresh start (s_time = 1) at the production time
    step, over the ramp's 40 tau (init_force_switch_on.f90:15-26,
    fluid_fluid.f90:42-49).  Charged systems run the ramp under the Langevin
    thermostat (THERMOSTAT 1): while the clamp softens the LJ cores, opposite
    charges fall together and the released Coulomb energy would heat the
    dilute channel (the DPD thermostat barely acts there) until beads cross
    the walls.  Returns (production params, state dict)."""
    p, r, v, t = build(inputs, kind, min_dist=min_dist)
    w = lib.Inputs.from_buffer_copy(inputs)
    w.s_time = 1
    w.force_switch_on = 1
    if any(q != 0.0 for q in w.charge):
        w.thermostat = 1
    wp = lib.derive(w)
    with lib.Context(wp, precision) as c:
        c.set_state(r, v, np.zeros_like(r), t)
        c.run_steps(1, ramp_steps(wp), want_obs=False)
        st = c.get_state()
    return p, {"r": st["r"], "v": st["v"], "a": st["a"], "mic": st["mic"], "t": t}


def small_cases():
    """Parity cases the oracle finishes in seconds."""
    return {
        "brush_small": (1, brush(8, 10, 300, 6.0, 6.0, 8.0, x_space=0.5, seed=11)),
        "melt_small": (2, melt(30, 20, 0.85, lj_rc=2.5, seed=12, n_cell=16)),
        "melt_wca": (2, melt(40, 10, 0.85, lj_rc=0.0, dt=0.005, seed=13, n_cell=16)),
        "melt_switch_on": (2, melt(40, 10, 0.85, lj_rc=0.0, dt=0.005, seed=14, n_cell=16,
                                   switch_on=True)),
        # stiff brush: harmonic bending of the grafted chains (BENDING)
        "brush_bend": (1, stiff(brush(8, 10, 300, 6.0, 6.0, 8.0, x_space=0.5, seed=17), 5.0, 0.4)),
        # charged brush, the reference's own Ewald form (compat) and the
        # corrected one with a real-space cut-off of 3
        "pe_brush_compat": (3, pe_brush(8, 12, 60, 6.0, 6.0, 8.0, x_space=0.5, n_kmax=4,
                                        compat=True, seed=15)),
        "pe_brush_small": (3, pe_brush(8, 12, 60, 12.0, 12.0, 14.0, x_space=0.5, n_kmax=4,
                                       rc_coul=3.0, seed=16)),
    }


THETA_NAMES = ("epsilon_LJ", "sigma", "k_FENE", "R0", "gamma_DPD")


def theta_of(inputs):
    """The parameters the adjoint differentiates: the common non-zero
    epsilon, the common sigma, FENE k and R0, and friction(1)."""
    eps = sorted({inputs.epsil[t][u] for t in range(NT) for u in range(NT)} - {0.0})
    sig = sorted({inputs.sigma_ii[t] for t in range(NT)})
    if len(eps) != 1 or len(sig) != 1:
        raise ValueError("epsilon and sigma must each take one value")
    return np.array([eps[0], sig[0], inputs.k_chain, inputs.r_chain, inputs.friction[0]])


def with_theta(inputs, theta):
    """A copy of `inputs` with the parameters set to theta (epsilon replaces
    every non-zero epsil entry, sigma every sigma_ii)."""
    x = lib.Inputs.from_buffer_copy(inputs)
    eps, sig, k, r0, gam = (float(v) for v in theta)
    for t in range(NT):
        x.sigma_ii[t] = sig
        for u in range(NT):
            if inputs.epsil[t][u] != 0.0:
                x.epsil[t][u] = eps
    x.k_chain = k
    x.r_chain = r0
    x.friction[0] = gam
    return x


def warmup_inputs(inputs, dt=0.002):
    """Start-up form of a melt: the same system with the purely repulsive
    WCA cut-off and FORCE_SWITCH_ON on a fresh start (s_time = 1), at a
    smaller time step.  The switch-on clamp r >= 1.11 sigma
    (fluid_fluid.f90:96) would remove the core but keep the attraction of an
    rc = 2.5 potential and collapse the melt, and at dt = 0.01 the re-expanding
    bonds heat the system until it blows up; relaxed under WCA at dt = 0.002
    over the 40 tau ramp, the melt then runs at the production cut-off and
    time step at T = 1."""
    w = lib.Inputs.from_buffer_copy(inputs)
    w.lj_rc = 0.0
    w.s_time = 1
    w.force_switch_on = 1
    w.dt = dt
    return w


def equilibrate(inputs, kind, min_dist=0.8, cool_steps=2000):
    """Seeded start, then on the GPU (melts only): the switch-on warm-up
    under WCA, and `cool_steps` of the production potential at dt = 0.002,
    whose DPD range (rc = 2.5) brings the warm-up's T ~ 5 back to kT before
    the production time step.  Returns (params, state dict: r, v, a, mic, t)."""
    p, r, v, t = build(inputs, kind, min_dist=min_dist)
    a = np.zeros_like(r)
    mic = np.zeros((p.n_part, 3), dtype=np.int32)
    if kind == 2:
        wi = warmup_inputs(inputs)
        wp = lib.derive(wi)
        with lib.Context(wp, lib.MFA_MIXED) as c:
            c.set_state(r, v, a, t)
            c.run_steps(1, ramp_steps(wp), want_obs=False)
            st = c.get_state()
        ci = lib.Inputs.from_buffer_copy(inputs)
        ci.dt = wi.dt
        with lib.Context(lib.derive(ci), lib.MFA_MIXED) as c:
            c.set_state(st["r"], st["v"], st["a"], t, st["mic"])
            c.run_steps(1, cool_steps, want_obs=False)
            st = c.get_state()
        r, v, mic = st["r"], st["v"], st["mic"]
        a = st["a"]
    return p, {"r": r, "v": v, "a": a, "mic": mic, "t": t}


def ramp_steps(params):
    """Steps until the switch-on clamp is gone (r_time = step*dt >= fso_time)."""
    return int(params.fso_time / params.dt) + 1


def build(inputs, kind, seed=None, min_dist=0.8):
    """Derive constants and generate the seeded initial state."""
    p = lib.derive(inputs)
    r, v, t = lib.generate(p, kind, inputs.iseed if seed is None else seed,
                           head_z=1.0, bond=0.97, min_dist=min_dist)
    return p, r, v, t
