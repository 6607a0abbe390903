// ISDF interpolation points by density-weighted centroidal Voronoi
// tessellation, the Lloyd iteration of shared/cvt.F90:8-195:
//   assign every grid point to its nearest centroid          (cvt.F90:77-88)
//   centroid = rho-weighted mean of its points; a centroid whose weight is
//   below 1e-10 keeps its previous position                  (:92-103)
//   snap each centroid, in index order, to the nearest grid point not taken
//   by an earlier centroid                                   (:104-128)
//   stop when the summed displacement of the points is below tol (:131-135)
// The assignment and the weighted sums (O(ngrid n_intp) per iteration) run
// on the GPU; the snap (O(n_intp)) runs on the host through a dense lattice
// lookup table instead of the reference's O(n_intp^2 ngrid) duplicate scan.
//
// Coordinates are in lattice units (grid index triples; bohr = h x units).
// Once the centroids are snapped they are lattice points, so squared
// distances are exact integers and ties resolve the same way everywhere:
// the lowest centroid index wins the assignment; the snap searches cubic
// shells s = 0, 1, ... around the rounded centroid, each in (z, y, x) order,
// and keeps the first free grid point at the smallest distance.

#include "ngw_kernels.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace ngw {

namespace {

constexpr int CVT_CHUNK = 1024;  // centroids staged in LDS per pass

// owner[g] = index of the nearest centroid (lowest index on ties)
__global__ __launch_bounds__(256) void cvt_assign_kernel(const int* grid, int ngrid, const double* cen, int ncen,
                                                         int* owner)
{
    __shared__ double sc[3][CVT_CHUNK];
    const int g = blockIdx.x * 256 + threadIdx.x;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (g < ngrid) {
        gx = (double)grid[3 * g];
        gy = (double)grid[3 * g + 1];
        gz = (double)grid[3 * g + 2];
    }
    double best = 1e300;
    int bi = 0;
    for (int c0 = 0; c0 < ncen; c0 += CVT_CHUNK) {
        const int cn = min(CVT_CHUNK, ncen - c0);
        __syncthreads();
        for (int i = threadIdx.x; i < cn; i += 256) {
            sc[0][i] = cen[3 * (c0 + i)];
            sc[1][i] = cen[3 * (c0 + i) + 1];
            sc[2][i] = cen[3 * (c0 + i) + 2];
        }
        __syncthreads();
        for (int i = 0; i < cn; ++i) {
            const double dx = gx - sc[0][i], dy = gy - sc[1][i], dz = gz - sc[2][i];
            const double d = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
            if (d < best) {
                best = d;
                bi = c0 + i;
            }
        }
    }
    if (g < ngrid) owner[g] = bi;
}

// wsum[4 c + {0,1,2,3}] += rho, rho x, rho y, rho z over the points of c
__global__ void cvt_accumulate_kernel(const int* grid, const double* rho, const int* owner, int ngrid, double* wsum)
{
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ngrid; g += gridDim.x * blockDim.x) {
        const int c = owner[g];
        const double r = rho[g];
        unsafeAtomicAdd(wsum + 4 * c, r);
        unsafeAtomicAdd(wsum + 4 * c + 1, r * (double)grid[3 * g]);
        unsafeAtomicAdd(wsum + 4 * c + 2, r * (double)grid[3 * g + 1]);
        unsafeAtomicAdd(wsum + 4 * c + 3, r * (double)grid[3 * g + 2]);
    }
}

__global__ void zero_kernel(double* x, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = 0.0;
}

// Dense lookup of grid indices over the bounding box of the lattice.
struct Lattice {
    int lo[3], n[3];
    std::vector<int> idx;  // -1 where no grid point
    int at(long i, long j, long k) const
    {
        i -= lo[0];
        j -= lo[1];
        k -= lo[2];
        if (i < 0 || j < 0 || k < 0 || i >= n[0] || j >= n[1] || k >= n[2]) return -1;
        return idx[(size_t)i + (size_t)n[0] * ((size_t)j + (size_t)n[1] * (size_t)k)];
    }
};

// Nearest free grid point to centroid c (lattice units); on equal distances
// the first one in shell, then (z, y, x), order.
int snap_one(const Lattice& L, const std::vector<char>& taken, const double c[3])
{
    const long r[3] = {std::lround(c[0]), std::lround(c[1]), std::lround(c[2])};
    int best = -1;
    double bd = 1e300;
    for (long s = 0;; ++s) {
        // lattice points in the shell max(|di|, |dj|, |dk|) == s around r
        for (long dk = -s; dk <= s; ++dk)
            for (long dj = -s; dj <= s; ++dj)
                for (long di = -s; di <= s; ++di) {
                    if (std::max(std::labs(di), std::max(std::labs(dj), std::labs(dk))) != s) continue;
                    const int g = L.at(r[0] + di, r[1] + dj, r[2] + dk);
                    if (g < 0 || taken[g]) continue;
                    const double ex = (double)(r[0] + di) - c[0], ey = (double)(r[1] + dj) - c[1],
                                 ez = (double)(r[2] + dk) - c[2];
                    const double d = (ex * ex + ey * ey) + ez * ez;
                    if (d < bd) {
                        bd = d;
                        best = g;
                    }
                }
        // every point outside shell s is at least s + 0.5 from c per axis
        if (best >= 0 && (double)s + 0.5 > std::sqrt(bd)) return best;
        if (s > 4 * (long)(L.n[0] + L.n[1] + L.n[2])) return best;  // whole box searched
    }
}

}  // namespace

}  // namespace ngw

using namespace ngw;

extern "C" int ngw_cvt_points(ngw_context* ctx, int ngrid, const int* grid, const double* rho, int n_intp,
                              const double* init, double h, double tol, int max_iter, int* intp, int* iters)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (ngrid <= 0 || n_intp <= 0 || n_intp > ngrid || !grid || !rho || !init || !intp || max_iter < 1 || !(h > 0))
        return fail(NGW_ERR_ARG, "bad sizes or arrays");
    NGW_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    // lattice table on the host
    Lattice L;
    int hi[3];
    for (int k = 0; k < 3; ++k) {
        L.lo[k] = grid[k];
        hi[k] = grid[k];
    }
    for (int g = 0; g < ngrid; ++g)
        for (int k = 0; k < 3; ++k) {
            L.lo[k] = std::min(L.lo[k], grid[3 * g + k]);
            hi[k] = std::max(hi[k], grid[3 * g + k]);
        }
    for (int k = 0; k < 3; ++k) L.n[k] = hi[k] - L.lo[k] + 1;
    const size_t cells = (size_t)L.n[0] * L.n[1] * L.n[2];
    if (cells > ((size_t)1 << 31)) return fail(NGW_ERR_ARG, "grid bounding box too large");
    L.idx.assign(cells, -1);
    for (int g = 0; g < ngrid; ++g) {
        const size_t o = (size_t)(grid[3 * g] - L.lo[0]) +
                         (size_t)L.n[0] * ((size_t)(grid[3 * g + 1] - L.lo[1]) +
                                           (size_t)L.n[1] * (size_t)(grid[3 * g + 2] - L.lo[2]));
        if (L.idx[o] >= 0) return fail(NGW_ERR_ARG, "duplicate grid point");
        L.idx[o] = g;
    }
    DevBuf dgrid, drho, dcen, downer, dw;
    NGW_TRY(dgrid.ensure(sizeof(int) * 3 * (size_t)ngrid));
    NGW_TRY(drho.ensure(sizeof(double) * (size_t)ngrid));
    NGW_TRY(dcen.ensure(sizeof(double) * 3 * (size_t)n_intp));
    NGW_TRY(downer.ensure(sizeof(int) * (size_t)ngrid));
    NGW_TRY(dw.ensure(sizeof(double) * 4 * (size_t)n_intp));
    NGW_HIP(hipMemcpyAsync(dgrid.ptr, grid, sizeof(int) * 3 * (size_t)ngrid, hipMemcpyHostToDevice, s));
    NGW_HIP(hipMemcpyAsync(drho.ptr, rho, sizeof(double) * (size_t)ngrid, hipMemcpyHostToDevice, s));
    std::vector<double> cen(3 * (size_t)n_intp), old(3 * (size_t)n_intp), w(4 * (size_t)n_intp);
    for (size_t i = 0; i < cen.size(); ++i) cen[i] = init[i] / h;
    old = cen;
    std::vector<char> taken(ngrid);
    int it = 0;
    for (it = 1; it <= max_iter; ++it) {
        NGW_HIP(hipMemcpyAsync(dcen.ptr, cen.data(), sizeof(double) * cen.size(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(zero_kernel, dim3(64), dim3(256), 0, s, dw.as<double>(), 4L * n_intp);
        hipLaunchKernelGGL(cvt_assign_kernel, dim3((ngrid + 255) / 256), dim3(256), 0, s, dgrid.as<int>(), ngrid,
                           dcen.as<double>(), n_intp, downer.as<int>());
        hipLaunchKernelGGL(cvt_accumulate_kernel, dim3(std::min((ngrid + 255) / 256, 2048)), dim3(256), 0, s,
                           dgrid.as<int>(), drho.as<double>(), downer.as<int>(), ngrid, dw.as<double>());
        NGW_HIP(hipGetLastError());
        NGW_HIP(hipMemcpyAsync(w.data(), dw.ptr, sizeof(double) * w.size(), hipMemcpyDeviceToHost, s));
        NGW_HIP(hipStreamSynchronize(s));
        std::fill(taken.begin(), taken.end(), 0);
        double diff = 0.0;
        for (int c = 0; c < n_intp; ++c) {
            double cand[3];
            const double wt = w[4 * (size_t)c];
            for (int k = 0; k < 3; ++k) cand[k] = wt >= 1e-10 ? w[4 * (size_t)c + 1 + k] / wt : old[3 * (size_t)c + k];
            const int g = snap_one(L, taken, cand);
            if (g < 0) return fail(NGW_ERR_ARG, "more interpolation points than free grid points");
            taken[g] = 1;
            intp[c] = g;
            double d2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double nk = (double)grid[3 * (size_t)g + k];
                d2 += (nk - old[3 * (size_t)c + k]) * (nk - old[3 * (size_t)c + k]);
                cen[3 * (size_t)c + k] = nk;
            }
            diff += std::sqrt(d2) * h;
        }
        old = cen;
        if (diff < tol) break;
    }
    if (iters) *iters = std::min(it, max_iter);
    return NGW_OK;
}
