// The specialised pass kernels with nontemporal output stores: fast.hip
// compiled a second time, exporting only its kernel table (fast_registry_nt).
#define PFFT_FAST_NT 1
#include "fast.hip"
