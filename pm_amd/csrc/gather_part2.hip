// One kernel group of the gather (k_migrate instantiations for the straight ray without anti-aliasing
// traveltime path), compiled as a translation unit of its own so the groups
// build in parallel.  pstm_hip.hip holds the kernel and its launcher
// pstm_gather_part2; the main translation unit calls it through dispatch().
#define PSTM_GATHER_PART 2
#include "pstm_hip.hip"
