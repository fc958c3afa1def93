// The two QPI sums on gfx950.
//
// 1. Autocorrelation over a k-plane (SurfaceDOSkk's jdos / jsdos,
// src/fermiarc.f90:496-560):
//
//   out(iq1, iq2) = sum_{ik1, ik2} sum_f a_f(ik1, ik2) a_f(ik1q, ik2q)
//   ik1q - 1 = (ik1 - 1 + iq1 - 1) mod m1,   ik2q - 1 = (ik2 - 1 + iq2 - 1) mod m2
//
// The shifted factor only ever reads the m1 x m2 corner of the plane, and the
// unshifted index enters only through (ik - 1) mod m, so the plane is first
// folded onto that torus, E_f(c1, c2) = sum of a_f over ik1 - 1 = c1 (mod m1),
// ik2 - 1 = c2 (mod m2), and the sum becomes the cyclic correlation
//
//   C(p1, p2) = sum_{c1 < m1, c2 < m2} sum_f E_f(c1, c2) a_f((c1 + p1) mod m1, (c2 + p2) mod m2)
//
// with out(iq1, iq2) = C((iq1 - 1) mod m1, (iq2 - 1) mod m2).  The correlation
// kernel takes one p2 per workgroup and one p1 per thread: per c2 it stages
// the row E_f(:, c2) and the row a_f(:, (c2 + p2) mod m2) twice over (so that
// c1 + p1 needs no wrap) in LDS; every thread then reads the E row at one
// address (broadcast) and the a row at consecutive addresses.  The sum is
// O((m1 m2)^2) FP64 multiply-adds, milliseconds at a 257 x 257 plane.
//
// 2. Linear autocorrelation along a k path, per energy (surfstat_jdos,
// src/surfstat.f90:891-907): with h = (nk - 1)/2 and s = iq - h,
//
//   out(iq, j) = sum_{ik = max(1, 1 - s)}^{min(2h + 1, 2h + 1 - s)} a(ik, j) a(ik + s, j)
//
// one thread per (iq, j, field), the terms added in the reference's order.
#include <algorithm>

#include "wtsurf_device.h"
#include "wtsurf_internal.h"

namespace wtsurf {

namespace {

constexpr int kCorrBlock = 256;

// E_f(c1, c2) for c1 < m1, c2 < m2; fields [nfield][nky][nkx], E [nfield][m2][m1].
__global__ void __launch_bounds__(kCorrBlock) fold_kernel(const double *__restrict__ a, int nkx,
                                                          int nky, int m1, int m2, int nfield,
                                                          double *__restrict__ e) {
    const long total = (long)nfield * m1 * m2;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int c1 = (int)(t % m1);
        const int c2 = (int)((t / m1) % m2);
        const int f = (int)(t / ((long)m1 * m2));
        const double *af = a + (size_t)f * nkx * nky;
        double s = 0.0;
        for (int i2 = c2; i2 < nky; i2 += m2)
            for (int i1 = c1; i1 < nkx; i1 += m1) s += af[i1 + (size_t)nkx * i2];
        e[t] = s;
    }
}

// C(p1, p2): block (p2, p1 tile); dynamic LDS 3 m1 doubles.
__global__ void __launch_bounds__(kCorrBlock) correlate_kernel(const double *__restrict__ a,
                                                               const double *__restrict__ e,
                                                               int nkx, int nky, int m1, int m2,
                                                               int nfield, double *__restrict__ c) {
    extern __shared__ __align__(16) double lds[];
    double *erow = lds;        // [m1]
    double *arow = lds + m1;   // [2 m1]: a_f(j mod m1, r)
    const int p2 = blockIdx.x;
    const int p1 = blockIdx.y * kCorrBlock + threadIdx.x;
    double acc = 0.0;
    for (int c2 = 0; c2 < m2; ++c2) {
        int r = c2 + p2;
        if (r >= m2) r -= m2;
        for (int f = 0; f < nfield; ++f) {
            const double *ef = e + ((size_t)f * m2 + c2) * m1;
            const double *af = a + (size_t)f * nkx * nky + (size_t)nkx * r;
            __syncthreads();  // the previous rows are no longer read
            for (int j = threadIdx.x; j < m1; j += kCorrBlock) {
                erow[j] = ef[j];
                const double v = af[j];
                arow[j] = v;
                arow[j + m1] = v;
            }
            __syncthreads();
            if (p1 < m1) {
                const double *ar = arow + p1;
                double s0 = 0.0, s1 = 0.0;
                int c1 = 0;
                for (; c1 + 1 < m1; c1 += 2) {
                    s0 = fma(erow[c1], ar[c1], s0);
                    s1 = fma(erow[c1 + 1], ar[c1 + 1], s1);
                }
                if (c1 < m1) s0 = fma(erow[c1], ar[c1], s0);
                acc += s0 + s1;
            }
        }
    }
    if (p1 < m1) c[(size_t)p2 * m1 + p1] = acc;
}

// out(iq1, iq2) = C((iq1 - 1) mod m1, (iq2 - 1) mod m2); out [nky][nkx].
__global__ void __launch_bounds__(kCorrBlock) expand_kernel(const double *__restrict__ c, int nkx,
                                                            int nky, int m1, int m2,
                                                            double *__restrict__ out) {
    const long total = (long)nkx * nky;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int i1 = (int)(t % nkx), i2 = (int)(t / nkx);
        out[t] = c[(size_t)(i2 % m2) * m1 + (i1 % m1)];
    }
}

// out_f(iq, j), fields and out [nfield][nomega][nk]; thread iq of row (j, f).
__global__ void __launch_bounds__(kCorrBlock) kpath_jdos_kernel(const double *__restrict__ a,
                                                                int nk, double *__restrict__ out) {
    const int iq = blockIdx.x * kCorrBlock + threadIdx.x;  // 0-based
    if (iq >= nk) return;
    const size_t row = (size_t)blockIdx.y * nk;
    const double *r = a + row;
    const int h = (nk - 1) / 2;
    const int sft = iq + 1 - h;
    const int lo = sft < 0 ? -sft : 0;                       // 0-based first ik
    const int hi = sft > 0 ? 2 * h - sft : 2 * h;            // 0-based last ik
    double s = 0.0;
    for (int ik = lo; ik <= hi; ++ik) s += r[ik] * r[ik + sft];
    out[row + iq] = s;
}

}  // namespace

int kplane_corr_max_m() { return kKplaneMaxM; }

int kpath_jdos_launch(const double *fields, int nk, int nomega, int nfield, double *out,
                      hipStream_t s) {
    if (nk < 1 || nomega < 1 || nfield < 1 || (long)nomega * nfield > 65535L * 1024) return -1;
    const long rows = (long)nomega * nfield;
    for (long r0 = 0; r0 < rows; r0 += 65535) {
        const int nr = (int)std::min<long>(rows - r0, 65535);
        hipLaunchKernelGGL(kpath_jdos_kernel, dim3((nk + kCorrBlock - 1) / kCorrBlock, nr),
                           dim3(kCorrBlock), 0, s, fields + (size_t)r0 * nk, nk,
                           out + (size_t)r0 * nk);
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int kplane_corr_launch(const KplaneCorrArgs &k, hipStream_t s) {
    if (k.nkx < 1 || k.nky < 1 || k.m1 < 1 || k.m2 < 1 || k.m1 > k.nkx || k.m2 > k.nky ||
        k.m1 > kKplaneMaxM || k.nfield < 1)
        return -1;
    const long nfold = (long)k.nfield * k.m1 * k.m2;
    const int gfold = (int)std::min<long>((nfold + kCorrBlock - 1) / kCorrBlock, 4096);
    hipLaunchKernelGGL(fold_kernel, dim3(gfold), dim3(kCorrBlock), 0, s, k.fields, k.nkx, k.nky,
                       k.m1, k.m2, k.nfield, k.folded);
    const dim3 grid(k.m2, (k.m1 + kCorrBlock - 1) / kCorrBlock);
    const size_t lds = (size_t)3 * k.m1 * sizeof(double);
    hipLaunchKernelGGL(correlate_kernel, grid, dim3(kCorrBlock), lds, s, k.fields, k.folded, k.nkx,
                       k.nky, k.m1, k.m2, k.nfield, k.corr);
    const long nout = (long)k.nkx * k.nky;
    const int gout = (int)std::min<long>((nout + kCorrBlock - 1) / kCorrBlock, 4096);
    hipLaunchKernelGGL(expand_kernel, dim3(gout), dim3(kCorrBlock), 0, s, k.corr, k.nkx, k.nky,
                       k.m1, k.m2, k.out);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace wtsurf
