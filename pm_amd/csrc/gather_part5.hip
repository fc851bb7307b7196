// One kernel group of the gather (k_migrate instantiations for the VTI with the learned traveltime network
// traveltime path), compiled as a translation unit of its own so the groups
// build in parallel.  pstm_hip.hip holds the kernel and its launcher
// pstm_gather_part5; the main translation unit calls it through dispatch().
#define PSTM_GATHER_PART 5
#include "pstm_hip.hip"
