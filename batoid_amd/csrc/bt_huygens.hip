// bt_huygens.hip -- Huygens PSF as a complex GEMM on MFMA (placeholder for
// the MFMA kernels; the FP64 direct sum lives in bt_field.hip).
#include <hip/hip_runtime.h>
#include "bt_common.h"

extern "C" int64_t bt_huygens_direct_workspace_bytes(int64_t n, int nx, int ny);

extern "C" int64_t bt_huygens_workspace_bytes(int64_t n, int nx, int ny)
{
    return bt_huygens_direct_workspace_bytes(n, nx, ny);
}

extern "C" int bt_huygens_gemm(bt_rays rays, const double* p0, const double* a1,
                               const double* a2, int nx, int ny, double t,
                               int ignore_vignetted, int precision, double* amp_out,
                               void* workspace, int64_t workspace_bytes, void* stream)
{
    (void)rays; (void)p0; (void)a1; (void)a2; (void)nx; (void)ny; (void)t;
    (void)ignore_vignetted; (void)precision; (void)amp_out; (void)workspace;
    (void)workspace_bytes; (void)stream;
    return bt::fail(BT_E_UNSUPPORTED, "bt_huygens_gemm: not built yet");
}
