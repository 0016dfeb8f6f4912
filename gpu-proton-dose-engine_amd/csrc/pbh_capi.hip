// pbh_capi.hip -- error strings and version of the C ABI (pencil_beam.cu:686-699).
#include "pbh_common.h"

extern "C" const char* pbh_strerror(int code)
{
    switch (code) {
    case PBH_OK: return "ok";
    case PBH_ERR_NULL_ARG: return "null argument";
    case PBH_ERR_BAD_DIM: return "invalid dimension / spacing";
    case PBH_ERR_ZGRID_MISMATCH: return "z grid mismatch";
    case PBH_ERR_HIP: return "HIP runtime error";
    case PBH_ERR_UNSUPPORTED: return "unsupported configuration";
    case PBH_ERR_NO_SMEM: return "insufficient LDS per workgroup";
    case PBH_ERR_WORKSPACE: return "workspace too small (query the matching *_workspace_bytes)";
    case PBH_ERR_SCRATCH: return "problem needs more than the fixed scratch of pbh_forward_noalloc "
                                 "(PBH_NOALLOC_SCRATCH_BYTES)";
    default: return "unknown error";
    }
}

extern "C" const char* pbh_version(void) { return "pbdose_hip 0.1 gfx950"; }
