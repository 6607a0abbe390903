// ISDF Coulomb kernel M = (ntrans/h^3) zeta^T V zeta on the device.
//
// V zeta_mu is the isolated-system Poisson solve of NanoGW's Zpoisson
// (shared/fft_module.F90z:122-177): put the column into the FFT box at the
// integer grid coordinates, R2C FFT, multiply by the truncated Coulomb kernel
// of Zcreate_coul_0D (:628-691), C2R FFT, pick the grid points and scale by
// 1/N.  Columns are processed in batches: one hand-written scatter kernel,
// one batched hipFFT D2Z, one kernel-multiply kernel, one batched Z2D, one
// gather kernel.  V zeta is kept for a group of batches (all columns when
// they fit in 16 GiB) and M(:, group) = zeta^T (V zeta_group) is one FP64
// MFMA GEMM per group (isdf_parallel_sym_UltraLowMem.F90:865-867): a GEMM
// per FFT batch (33 columns at 158^3) would give too few tiles to fill the
// GPU.

#include <hipfft/hipfft.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "ngw_kernels.h"

namespace ngw {

#define NGW_FFT(call)                                                                        \
    do {                                                                                     \
        hipfftResult r_ = (call);                                                            \
        if (r_ != HIPFFT_SUCCESS)                                                            \
            return ::ngw::fail(NGW_ERR_HIP, std::string(#call " -> hipfft error ") + std::to_string((int)r_)); \
    } while (0)

// Box offset of each grid point (x fastest), periodic wrap of the centred
// coordinates as in Zpoisson's fft_box%Zbox(-n/2:n/2-1, ...) layout.
__global__ void box_offsets_kernel(const int* grid, int ngrid, int n1, int n2, int n3, long* off)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ngrid) return;
    int x = grid[3 * i], y = grid[3 * i + 1], z = grid[3 * i + 2];
    x = ((x % n1) + n1) % n1;
    y = ((y % n2) + n2) % n2;
    z = ((z % n3) + n3) % n3;
    off[i] = ((long)z * n2 + y) * n1 + x;
}

// box[b][off[i]] = zeta(i, col0 + b); the rest of the box is zero.
__global__ void scatter_kernel(const double* zeta, long ldz, int ngrid, const long* off, long boxsz,
                               double* box, int nb)
{
    const int b = blockIdx.y;
    const double* zc = zeta + (long)b * ldz;
    double* bx = box + (long)b * boxsz;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ngrid; i += gridDim.x * blockDim.x)
        bx[off[i]] = zc[i];
}

// V(G) box multiply: fbox[b][g] *= vcoul[g]
__global__ void kernel_multiply(double2* fbox, const double* vcoul, long csz, int nb)
{
    const long total = csz * nb;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const double v = vcoul[t % csz];
        double2 f = fbox[t];
        f.x *= v;
        f.y *= v;
        fbox[t] = f;
    }
}

// vz(i, b) = box[b][off[i]] * scale
__global__ void gather_kernel(const double* box, long boxsz, const long* off, int ngrid, double scale,
                              double* vz, long ldv)
{
    const int b = blockIdx.y;
    const double* bx = box + (long)b * boxsz;
    double* vc = vz + (long)b * ldv;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ngrid; i += gridDim.x * blockDim.x)
        vc[i] = bx[off[i]] * scale;
}

// v(G) = 8 pi (1 - cos(|G| R_c)) / |G|^2,  v(0) = 8 pi R_c^2 / 2   (trunc > 0)
//      = 8 pi / |G|^2, v(0) = 0                                     (trunc <= 0)
// on the R2C half box, G_i = 2 pi j_i / (n_i h).
__global__ void coulomb_kernel(int n1, int n2, int n3, double h, double trunc, double* vcoul)
{
    const int h1 = n1 / 2 + 1;
    const long csz = (long)h1 * n2 * n3;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < csz; t += (long)gridDim.x * blockDim.x) {
        int i1 = (int)(t % h1);
        int i2 = (int)((t / h1) % n2);
        int i3 = (int)(t / ((long)h1 * n2));
        int j1 = i1;  // 0..n1/2
        int j2 = i2 < n2 / 2 ? i2 : i2 - n2;
        int j3 = i3 < n3 / 2 ? i3 : i3 - n3;
        const double pi = 3.1415926535897932384626433832795;
        double g1 = 2.0 * pi / (n1 * h) * j1, g2 = 2.0 * pi / (n2 * h) * j2, g3 = 2.0 * pi / (n3 * h) * j3;
        double ekin = g1 * g1 + g2 * g2 + g3 * g3;
        const double qzero = 1.0e-6;
        double v;
        if (trunc > 0.0)
            v = (ekin < qzero * qzero) ? trunc * trunc / 2.0 : (1.0 - cos(trunc * sqrt(ekin))) / ekin;
        else
            v = (ekin < qzero * qzero) ? 0.0 : 1.0 / ekin;
        vcoul[t] = 8.0 * pi * v;
    }
}

}  // namespace ngw

using namespace ngw;

// zeta_dev: zeta is a device array (ld_zeta) used in place; otherwise a host
// array copied to the device.  M is returned to the host.
static int coulomb_impl(ngw_context* ctx, int ngrid, int n_intp, const double* zeta, long ld_zeta, bool zeta_dev,
                        const int* grid, const int* nfft, double h, double trunc, double ntrans, double* mmtrx,
                        int ld_m)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (ngrid <= 0 || n_intp <= 0 || ld_zeta < ngrid || ld_m < n_intp || !zeta || !grid || !nfft || !mmtrx)
        return fail(NGW_ERR_ARG, "bad sizes or null arrays");
    const int n1 = nfft[0], n2 = nfft[1], n3 = nfft[2];
    if (n1 <= 1 || n2 <= 1 || n3 <= 1 || (n1 % 2) || (n2 % 2) || (n3 % 2))
        return fail(NGW_ERR_ARG, "nfft must be even and > 1");
    if (!(h > 0.0)) return fail(NGW_ERR_ARG, "h must be positive");
    for (int i = 0; i < ngrid; ++i)
        for (int k = 0; k < 3; ++k) {
            int n = nfft[k], g = grid[3 * i + k];
            if (g < -n / 2 || g >= n / 2) return fail(NGW_ERR_ARG, "grid coordinate outside the FFT box");
        }
    NGW_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const long boxsz = (long)n1 * n2 * n3;
    const long csz = (long)(n1 / 2 + 1) * n2 * n3;
    // batch so that the real and complex boxes stay near 2 GiB
    int nb = (int)std::max(1L, std::min((long)n_intp, (2L << 30) / (8 * boxsz + 16 * csz)));
    nb = std::min(nb, 256);
    // columns of V zeta kept for one GEMM: a multiple of nb, up to 16 GiB
    const long vz_cols = std::max(1L, (16L << 30) / (8L * ngrid));
    const int grp = (int)std::min((long)n_intp, std::max((long)nb, vz_cols / nb * nb));
    DevBuf dzeta, dvz, dbox, dfbox, dvc, doff, dgrid, dm;
    struct Guard {
        DevBuf* b[8];
        ~Guard() { for (DevBuf* x : b) x->release(); }
    } guard{{&dzeta, &dvz, &dbox, &dfbox, &dvc, &doff, &dgrid, &dm}};
    if (!zeta_dev) NGW_TRY(dzeta.ensure(sizeof(double) * (size_t)ngrid * n_intp));
    NGW_TRY(dvz.ensure(sizeof(double) * (size_t)ngrid * grp));
    NGW_TRY(dbox.ensure(sizeof(double) * (size_t)boxsz * nb));
    NGW_TRY(dfbox.ensure(sizeof(double) * 2 * (size_t)csz * nb));
    NGW_TRY(dvc.ensure(sizeof(double) * (size_t)csz));
    NGW_TRY(doff.ensure(sizeof(long) * (size_t)ngrid));
    NGW_TRY(dgrid.ensure(sizeof(int) * 3 * (size_t)ngrid));
    NGW_TRY(dm.ensure(sizeof(double) * (size_t)n_intp * n_intp));
    const double* zd = zeta;  // device zeta and its leading dimension
    long ldz = ld_zeta;
    if (!zeta_dev) {
        if (ld_zeta == ngrid) {
            NGW_HIP(hipMemcpyAsync(dzeta.ptr, zeta, sizeof(double) * (size_t)ngrid * n_intp, hipMemcpyHostToDevice,
                                   s));
        } else {
            NGW_HIP(hipMemcpy2DAsync(dzeta.ptr, sizeof(double) * ngrid, zeta, sizeof(double) * ld_zeta,
                                     sizeof(double) * ngrid, n_intp, hipMemcpyHostToDevice, s));
        }
        zd = dzeta.as<double>();
        ldz = ngrid;
    }
    NGW_HIP(hipMemcpyAsync(dgrid.ptr, grid, sizeof(int) * 3 * (size_t)ngrid, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(box_offsets_kernel, dim3((ngrid + 255) / 256), dim3(256), 0, s, dgrid.as<int>(), ngrid,
                       n1, n2, n3, doff.as<long>());
    NGW_HIP(hipGetLastError());
    hipLaunchKernelGGL(coulomb_kernel, dim3(1024), dim3(256), 0, s, n1, n2, n3, h, trunc, dvc.as<double>());
    NGW_HIP(hipGetLastError());
    hipfftHandle pf = 0, pb = 0;
    int dims[3] = {n3, n2, n1};
    int cdims[3] = {n3, n2, n1 / 2 + 1};
    struct PlanGuard {
        hipfftHandle* p[2];
        ~PlanGuard() { for (auto* x : p) if (*x) hipfftDestroy(*x); }
    } pg{{&pf, &pb}};
    NGW_FFT(hipfftPlanMany(&pf, 3, dims, dims, 1, (int)boxsz, cdims, 1, (int)csz, HIPFFT_D2Z, nb));
    NGW_FFT(hipfftPlanMany(&pb, 3, dims, cdims, 1, (int)csz, dims, 1, (int)boxsz, HIPFFT_Z2D, nb));
    NGW_FFT(hipfftSetStream(pf, s));
    NGW_FFT(hipfftSetStream(pb, s));
    const double scale = 1.0 / ((double)n1 * (double)n2 * (double)n3);
    const double norm = ntrans / (h * h * h);
    int g0 = 0;  // first column of the current group
    for (int c0 = 0; c0 < n_intp; c0 += nb) {
        const int cb = std::min(nb, n_intp - c0);
        double* vz = dvz.as<double>() + (size_t)(c0 - g0) * ngrid;
        NGW_HIP(hipMemsetAsync(dbox.ptr, 0, sizeof(double) * (size_t)boxsz * nb, s));
        hipLaunchKernelGGL(scatter_kernel, dim3(std::min((ngrid + 255) / 256, 1024), cb), dim3(256), 0, s,
                           zd + (size_t)c0 * ldz, ldz, ngrid, doff.as<long>(), boxsz,
                           dbox.as<double>(), cb);
        NGW_HIP(hipGetLastError());
        NGW_FFT(hipfftExecD2Z(pf, dbox.as<double>(), reinterpret_cast<hipfftDoubleComplex*>(dfbox.ptr)));
        hipLaunchKernelGGL(kernel_multiply, dim3(4096), dim3(256), 0, s, dfbox.as<double2>(), dvc.as<double>(),
                           csz, nb);
        NGW_HIP(hipGetLastError());
        NGW_FFT(hipfftExecZ2D(pb, reinterpret_cast<hipfftDoubleComplex*>(dfbox.ptr), dbox.as<double>()));
        hipLaunchKernelGGL(gather_kernel, dim3(std::min((ngrid + 255) / 256, 1024), cb), dim3(256), 0, s,
                           dbox.as<double>(), boxsz, doff.as<long>(), ngrid, scale, vz, (long)ngrid);
        NGW_HIP(hipGetLastError());
        const int gend = c0 + cb;
        if (gend - g0 == grp || gend == n_intp) {
            // M(:, g0:gend) = norm * zeta^T vz
            NGW_TRY(gemm_f64_tn(s, n_intp, gend - g0, ngrid, norm, zd, ldz, dvz.as<double>(), ngrid,
                                dm.as<double>() + (size_t)g0 * n_intp, n_intp));
            g0 = gend;
        }
    }
    std::vector<double> hm((size_t)n_intp * n_intp);
    NGW_HIP(hipMemcpyAsync(hm.data(), dm.ptr, sizeof(double) * hm.size(), hipMemcpyDeviceToHost, s));
    NGW_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < n_intp; ++j)
        std::memcpy(mmtrx + (size_t)j * ld_m, &hm[(size_t)j * n_intp], sizeof(double) * n_intp);
    return NGW_OK;
}

extern "C" int ngw_isdf_coulomb_kernel(ngw_context* ctx, int ngrid, int n_intp, const double* zeta,
                                       long ld_zeta, const int* grid, const int* nfft, double h,
                                       double trunc, double ntrans, double* mmtrx, int ld_m)
{
    return coulomb_impl(ctx, ngrid, n_intp, zeta, ld_zeta, false, grid, nfft, h, trunc, ntrans, mmtrx, ld_m);
}

extern "C" int ngw_isdf_coulomb_kernel_dev(ngw_context* ctx, int ngrid, int n_intp, const double* d_zeta,
                                           long ld_zeta, const int* grid, const int* nfft, double h,
                                           double trunc, double ntrans, double* mmtrx, int ld_m)
{
    return coulomb_impl(ctx, ngrid, n_intp, d_zeta, ld_zeta, true, grid, nfft, h, trunc, ntrans, mmtrx, ld_m);
}
