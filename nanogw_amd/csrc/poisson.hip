// ISDF Coulomb kernel: batched FFT Poisson solves and zeta^T V zeta.
#include "ngw_kernels.h"

extern "C" int ngw_isdf_coulomb_kernel(ngw_context* ctx, int ngrid, int n_intp, const double* zeta,
                                       long ld_zeta, const int* grid, const int* nfft, double h,
                                       double trunc, double ntrans, double* mmtrx, int ld_m)
{
    (void)ctx; (void)ngrid; (void)n_intp; (void)zeta; (void)ld_zeta; (void)grid; (void)nfft;
    (void)h; (void)trunc; (void)ntrans; (void)mmtrx; (void)ld_m;
    return ngw::fail(NGW_ERR_STATE, "not implemented yet");
}
