// One kernel group of the gather (k_migrate instantiations for the curved-ray
// traveltime read from the learned network's t0 lattice), compiled as a
// translation unit of its own so the groups build in parallel.  pstm_hip.hip
// holds the kernel and its launcher pstm_gather_part7; the main translation
// unit calls it through dispatch().
#define PSTM_GATHER_PART 7
#include "pstm_hip.hip"
