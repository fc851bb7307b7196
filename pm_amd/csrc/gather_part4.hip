// One kernel group of the gather (k_migrate instantiations for the VTI
// traveltime path), compiled as a translation unit of its own so the groups
// build in parallel.  pstm_hip.hip holds the kernel and its launcher
// pstm_gather_part4; the main translation unit calls it through dispatch().
#define PSTM_GATHER_PART 4
#include "pstm_hip.hip"
