// pbh_mixed.hip -- placeholder for the mixed-precision forward forms.
#include "pbh_common.h"
namespace pbh { struct FwdArgs; }
using namespace pbh;
int pbh_forward_mixed(const FwdArgs&, int, void*, size_t, hipStream_t) { return PBH_ERR_UNSUPPORTED; }
size_t pbh_forward_mixed_ws(const PbhGrid*, const PbhLattice*, int) { return 0; }
