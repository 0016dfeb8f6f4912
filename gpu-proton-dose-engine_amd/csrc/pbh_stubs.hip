// temporary: entry points not yet implemented
#include "pbh_common.h"
extern "C" size_t pbh_gamma_workspace_bytes(const PbhGrid*) { return 0; }
extern "C" int pbh_gamma_index(const PbhGrid*, const float*, const float*, double, double, double,
                    double, float*, unsigned long long*, void*, size_t, void*) { return PBH_ERR_UNSUPPORTED; }
