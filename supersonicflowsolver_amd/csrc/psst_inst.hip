// psst_inst.hip -- instantiates the solver's launch templates (and through
// them every kernel) for one (dimension, state type) pair, given on the
// command line as PSST_INST_DIM (2 or 3) and PSST_INST_T (double or float).
// The build compiles this file four times, in parallel.
#include "psst_device.hpp"

#if !defined(PSST_INST_DIM) || !defined(PSST_INST_T)
#error "PSST_INST_DIM and PSST_INST_T must be defined"
#endif

PSST_DIM_T_ENTRY_POINTS(, PSST_INST_DIM, PSST_INST_T)
