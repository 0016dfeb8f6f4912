// temporary: entry points not yet implemented
#include "pbh_common.h"
extern "C" size_t pbh_adjoint_workspace_bytes(const PbhGrid*, const PbhLattice*, int) { return 0; }
extern "C" int pbh_adjoint(const PbhGrid*, const PbhLattice*, const float*, const double*, const double*,
                const float*, float*, double*, double*, int, void*, size_t, void*) { return PBH_ERR_UNSUPPORTED; }
extern "C" size_t pbh_objective_workspace_bytes(const PbhGrid*) { return 0; }
extern "C" int pbh_objective(const PbhGrid*, const float*, const float*, const float*, float*, double*,
                  void*, size_t, void*) { return PBH_ERR_UNSUPPORTED; }
extern "C" size_t pbh_gamma_workspace_bytes(const PbhGrid*) { return 0; }
extern "C" int pbh_gamma_index(const PbhGrid*, const float*, const float*, double, double, double,
                    double, float*, unsigned long long*, void*, size_t, void*) { return PBH_ERR_UNSUPPORTED; }
