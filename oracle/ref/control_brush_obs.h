/* control_simulation.h switches for the brush builds of the reference with
   the observation profiles (PROFILES: density and velocity profiles in
   observation.f90:336-350):
   brush grafted on the bottom wall in a 9-3 wall channel, DPD thermostat,
   explicit DPD cut-off 1.0, linked-cell binning. */
#define SYSTEM 3
#define SYMMETRY 0
#define WALL 2
#define SOLVENT 1
#define THERMOSTAT 0
#define DPD_WEIGHT 0
#define DPD_CUT_OFF 1.0
#define STORE 0
#define NO_WARNS
#define BIN_TYPE 0
#define RINGS 0
#define FIXCM 0
#define CHAIN_TYPE 0
#define CHAIN_BC 0
#define EXT_FORCE 0
#define PROFILES
