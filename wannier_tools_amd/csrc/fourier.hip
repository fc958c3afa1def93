// Principal-layer Fourier sum H00(k), H01(k) for a batch of k points.
//
// Restates ham_qlayer2qlayer (src/ham_qlayer2qlayer.f90:45-97):
//   Hij(d) = sum_{R : R3 = d} e^{i 2 pi (k1 R1 + k2 R2)} H(R)
//   H00 block (i, j)    = Hij(j - i),  i, j in 1..Np
//   H01 block (i, j-Np) = Hij(j - i),  j in Np+1..2Np
// only for |d| <= ijmax.  The R list arrives sorted by R3 (stable, so every
// Hij(d) sums its R in the file order the reference uses) and only the slots
// d in [-(Np-1), 2Np-1] that H00/H01 can reference are evaluated.
//
// Stage 1 tabulates the phases e^{i 2 pi k.R} for all (k, R) once.  Stage 2
// is one workgroup per (k, slot): each thread owns Hij(d) elements, sums over
// the slot's R and scatters the result to every H00/H01 block using that
// slot, writing the NP x NP padded layout the decimation kernel reads.
#include "wtsurf_device.h"
#include "wtsurf_internal.h"

namespace wtsurf {

// Pi as the reference truncates it (src/module.f90:719).
constexpr double kRefPi = 3.14159265358979;

__global__ void phase_kernel(const double *__restrict__ kpts, const int *__restrict__ irvec,
                             int nk, int nr, double2 *__restrict__ phase) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nk * nr) return;
    const int ik = (int)(idx / nr), ir = (int)(idx - (long)ik * nr);
    const double kdotr = kpts[2 * ik] * (double)irvec[3 * ir] +
                         kpts[2 * ik + 1] * (double)irvec[3 * ir + 1];
    const double ph = (2.0 * kRefPi) * kdotr;
    double s, c;
    sincos(ph, &s, &c);
    phase[idx] = make_double2(c, s);
}

__global__ void hij_scatter_kernel(FourierArgs a) {
    const int ik = blockIdx.x;
    const int slot = blockIdx.y;
    const int d = a.d_lo + slot;
    const int nw = a.nw, np = a.np, NP = a.NP;
    const int r0 = a.slot_begin[slot], r1 = a.slot_begin[slot + 1];
    const double2 *ph = a.phase + (size_t)ik * a.nr;
    double2 *H00 = a.hk + (size_t)ik * 2 * NP * NP;
    double2 *H01 = H00 + (size_t)NP * NP;
    const int nw2 = nw * nw;
    for (int e = threadIdx.x; e < nw2; e += blockDim.x) {
        double2 acc = make_double2(0.0, 0.0);
        for (int ir = r0; ir < r1; ++ir) {
            const double2 h = a.hr[(size_t)ir * nw2 + e];
            const double2 p = ph[ir];
            acc.x += h.x * p.x - h.y * p.y;
            acc.y += h.x * p.y + h.y * p.x;
        }
        const int rr = e % nw, cc = e / nw;
        for (int i = 1; i <= np; ++i) {
            const int j = i + d;  // block column in 1..2Np
            if (j < 1 || j > 2 * np) continue;
            double2 *dst = (j <= np) ? H00 : H01;
            const int jb = (j <= np) ? j : j - np;
            dst[(nw * (i - 1) + rr) + (size_t)(nw * (jb - 1) + cc) * NP] = acc;
        }
    }
}

int fourier_launch(const FourierArgs &a, hipStream_t s) {
    if (a.nk <= 0) return 0;
    const size_t hk_bytes = (size_t)a.nk * 2 * a.NP * a.NP * sizeof(double2);
    if (hipMemsetAsync(a.hk, 0, hk_bytes, s) != hipSuccess) return -1;
    const long nph = (long)a.nk * a.nr;
    if (nph > 0) {
        const int tpb = 256;
        hipLaunchKernelGGL(phase_kernel, dim3((unsigned)((nph + tpb - 1) / tpb)), dim3(tpb), 0, s,
                           a.kpts, a.irvec, a.nk, a.nr, a.phase);
    }
    if (a.nslot > 0) {
        hipLaunchKernelGGL(hij_scatter_kernel, dim3(a.nk, a.nslot), dim3(256), 0, s, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace wtsurf
