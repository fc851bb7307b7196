// One kernel group of the gather (k_migrate instantiations for the curved ray
// traveltime path), compiled as a translation unit of its own so the groups
// build in parallel.  pstm_hip.hip holds the kernel and its launcher
// pstm_gather_part3; the main translation unit calls it through dispatch().
#define PSTM_GATHER_PART 3
#include "pstm_hip.hip"
