/* control_simulation.h switches for the melt builds of the reference:
   periodic in three directions, DPD thermostat over the LJ range,
   linked-cell binning. */
#define SYSTEM 2
#define SYMMETRY 1
#define WALL 2
#define SOLVENT 1
#define THERMOSTAT 0
#define DPD_WEIGHT 0
#define STORE 0
#define NO_WARNS
#define BIN_TYPE 0
#define RINGS 0
#define FIXCM 0
#define CHAIN_TYPE 0
#define CHAIN_BC 0
#define EXT_FORCE 0
