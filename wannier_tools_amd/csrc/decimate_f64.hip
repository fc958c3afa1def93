// Batched Lopez-Sancho decimation in FP64 on gfx950.
//
// Restates surfgreen_1985 (src/surfgreen.f90:74-134) for many (k, omega)
// instances in one persistent launch.  One workgroup (8 waves) owns one
// instance at a time and pulls the next instance id from a device counter
// when its decimation converges, so uneven iteration counts do not idle the
// chip.  Per iteration:
//
//   g = (z - eps_i)^-1                       Gauss-Jordan, matrix in VGPRs
//   [m1 ; m2]        = [alpha ; beta] g      dual GEMM, shared operand g
//   [m1 b ; m2 b]    -> eps_i, eps_s += m1 b ; beta'  = m2 b   (b = beta)
//   [m2 a ; m1 a]    -> eps_i, eps_t += m2 a ; alpha' = m1 a   (a = alpha)
//   stop when sum |alpha'_ij| <= accuracy
//
// which is the reference's 1 inverse + 6 GEMMs grouped into three products
// that each stage one shared operand through LDS.  GEMMs run on the
// 16x16x4 FP64 MFMA (4 real products per complex tile step).  The three
// final inverses produce only the diagonals the traces need (or the full
// G matrices when gfull is set).
#include "wtsurf_device.h"
#include "wtsurf_internal.h"

namespace wtsurf {

enum { M_A0, M_A1, M_B0, M_B1, M_EI, M_ES, M_ET, M_M1, M_M2, M_G };

template <int Q>
struct DecimSmem {
    static constexpr int NP = 16 * Q;
    double2 as0[2][kChunk][NP];
    double2 as1[2][kChunk][NP];
    double2 bs[2][kChunk][NP + 1];  // +1: staging writes walk k fastest
    double2 colbuf[2][NP];
    double2 rowbuf[2][NP];
    double2 diag[NP];
    int pivrow[NP];
    int pivstep[NP];
    double red[kWaves];
    int cand;
    int inst;
};

// ---------------------------------------------------------------------------
// Gauss-Jordan inverse of X = z*1 - E (the first n x n block of the padded
// E), with partial pivoting on |re|+|im| as zgetrf does.  Thread
// (tr, tc) = (tid % 16, tid / 16), tid < 256, keeps X[tr+16a][tc+16b]
// (a, b < Q) in registers; pivot column and pivot row travel through LDS.  Rows are not
// swapped: the pivot of step k is row pivrow[k], and at the end
// X^-1[i][j] = S[pivrow[i]][pivstep[j]].
//
// out_mode 0: write X^-1 into the padded matrix G (pad stays untouched).
// out_mode 1: write X^-1 diagonal into sm.diag.
// out_mode 2: both sm.diag and the unpadded n x n matrix `full`.
// Returns true if an exactly zero pivot column was met; the result is then
// the identity, as the reference's inv returns (zgesv skips the solve and
// inv copies back the identity right-hand side, src/inverse.f90:44-55).
template <int Q>
__device__ bool gj_inverse(DecimSmem<Q> &sm, const double2 *__restrict__ E,
                           double2 z, int n, double2 *__restrict__ G,
                           double2 *__restrict__ full, int out_mode) {
    constexpr int NP = 16 * Q;
    const int tid = opaque_tid();
    // waves 0-3 hold the matrix; waves 4-7 only join the barriers
    const bool holder = uniform(tid >> 6) < kGJThreads / kWave;
    const int tr = tid & 15, tc = (tid >> 4) & 15;
    double2 S[Q][Q];
#pragma unroll
    for (int a = 0; a < Q; ++a)
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            const int i = tr + 16 * a, j = tc + 16 * b;
            double2 v = make_double2(0.0, 0.0);
            if (holder && i < n && j < n) {
                const double2 e = E[i + j * NP];
                v = make_double2((i == j ? z.x : 0.0) - e.x, (i == j ? z.y : 0.0) - e.y);
            }
            S[a][b] = v;
        }
    unsigned pmask = 0;  // bit a: row tr+16a already used as a pivot
    bool singular = false;

#pragma unroll
    for (int bk = 0; bk < Q; ++bk)
    for (int kl = 0; kl < 16; ++kl) {
        const int k = 16 * bk + kl;
        if (k >= n || singular) break;
        const int buf = k & 1;
        if (holder && tc == kl) {
            double best = -1.0;
            int bi = 0x7fffffff;
#pragma unroll
            for (int a = 0; a < Q; ++a) {
                const double2 v = S[a][bk];
                const int i = tr + 16 * a;
                sm.colbuf[buf][i] = v;
                const double c = fabs(v.x) + fabs(v.y);
                if (!((pmask >> a) & 1u) && i < n && c > best) {
                    best = c;
                    bi = i;
                }
            }
#pragma unroll
            for (int off = 8; off >= 1; off >>= 1) {
                const double ob = __shfl_xor(best, off, 16);
                const int oi = __shfl_xor(bi, off, 16);
                if (ob > best || (ob == best && oi < bi)) {
                    best = ob;
                    bi = oi;
                }
            }
            if (tr == 0) sm.cand = (best > 0.0) ? bi : -1;
        }
        __syncthreads();
        const int p = uniform(sm.cand);
        if (p < 0) {
            singular = true;
            continue;
        }
        const int ap = p >> 4;
        if (holder && tr == (p & 15)) {
            const double2 r = crecip(sm.colbuf[buf][p]);
            pmask |= 1u << ap;
#pragma unroll
            for (int a = 0; a < Q; ++a) {
                if (a != ap) continue;
#pragma unroll
                for (int b = 0; b < Q; ++b) {
                    const int j = tc + 16 * b;
                    const double2 v = (j == k) ? r : cmul(S[a][b], r);
                    S[a][b] = v;
                    sm.rowbuf[buf][j] = v;
                }
            }
        }
        if (tid == 0) {
            sm.pivrow[k] = p;
            sm.pivstep[p] = k;
        }
        __syncthreads();
        if (!holder) continue;
#pragma unroll
        for (int a = 0; a < Q; ++a) {
            const int i = tr + 16 * a;
            if (i == p) continue;
            const double2 f = sm.colbuf[buf][i];
#pragma unroll
            for (int b = 0; b < Q; ++b) {
                const int j = tc + 16 * b;
                const double2 rb = sm.rowbuf[buf][j];
                double2 s = (j == k) ? make_double2(0.0, 0.0) : S[a][b];
                s.x -= f.x * rb.x - f.y * rb.y;
                s.y -= f.x * rb.y + f.y * rb.x;
                S[a][b] = s;
            }
        }
    }
    __syncthreads();  // pivrow / pivstep complete
    if (!holder) {
    } else if (singular) {
        // identity result
#pragma unroll
        for (int a = 0; a < Q; ++a)
#pragma unroll
            for (int b = 0; b < Q; ++b) {
                const int i = tr + 16 * a, j = tc + 16 * b;
                if (i < n && j < n) {
                    const double2 v = make_double2(i == j ? 1.0 : 0.0, 0.0);
                    if (out_mode == 0) G[i + j * NP] = v;
                    if (out_mode == 2) full[i + j * n] = v;
                    if (out_mode >= 1 && i == j) sm.diag[i] = v;
                }
            }
    } else {
#pragma unroll
        for (int a = 0; a < Q; ++a)
#pragma unroll
            for (int b = 0; b < Q; ++b) {
                const int rho = tr + 16 * a, kap = tc + 16 * b;
                if (rho < n && kap < n) {
                    const int i = sm.pivstep[rho], j = sm.pivrow[kap];
                    if (out_mode == 0) G[i + j * NP] = S[a][b];
                    if (out_mode == 2) full[i + j * n] = S[a][b];
                    if (out_mode >= 1 && i == j) sm.diag[i] = S[a][b];
                }
            }
    }
    __syncthreads();
    return singular;
}

// ---------------------------------------------------------------------------
// Dual GEMM sharing the right operand:  P0 = A0 * B,  P1 = A1 * B  (padded
// NP x NP, K = kdim rounded up to the chunk).  B and the A column panels are
// staged through LDS in chunks of kChunk columns, double-buffered with a
// register prefetch of the next chunk.  Tile t in [0, 2*Q*Q) belongs to wave
// t % 8; accumulators stay in registers for the whole K loop.
//
//   MODE 0: O0 = P0, O1 = P1
//   MODE 1: E1 += P0, E2 += P0 (E1 first, as the reference adds), O1 = P1
//   MODE 2: as MODE 1, and returns sum_ij |P1_ij| to every thread
template <int Q, int MODE>
__device__ double dual_gemm(DecimSmem<Q> &sm, const double2 *__restrict__ A0,
                            const double2 *__restrict__ A1,
                            const double2 *__restrict__ B, int kdim,
                            double2 *__restrict__ O0, double2 *__restrict__ O1,
                            double2 *__restrict__ E1, double2 *__restrict__ E2) {
    constexpr int NP = 16 * Q;
    constexpr int TILES = 2 * Q * Q;
    constexpr int TPW = (TILES + kWaves - 1) / kWaves;
    const int tid = opaque_tid();
    const int wave = uniform(tid >> 6), lane = tid & 63;
    const int lr = lane & 15, lq = lane >> 4;

    v4d accR[TPW], accI[TPW];
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
        accR[s] = v4d{0.0, 0.0, 0.0, 0.0};
        accI[s] = v4d{0.0, 0.0, 0.0, 0.0};
    }

    const int nchunks = (kdim + kChunk - 1) / kChunk;
    // Staging: element e = tid (+ kBlock) of the B rows and the A0/A1 column
    // panels of a chunk, loaded one chunk ahead into registers.
    constexpr int PER = kChunk * NP;
    static_assert(PER <= 2 * kBlock, "chunk larger than the register stage");
    const int e0 = tid, e1 = tid + kBlock;
    const bool has0 = e0 < PER, has1 = e1 < PER;
    const int bo0 = e0 % kChunk + (e0 / kChunk) * NP, ao0 = e0 % NP + (e0 / NP) * NP;
    const int bo1 = e1 % kChunk + (e1 / kChunk) * NP, ao1 = e1 % NP + (e1 / NP) * NP;
    double2 b0, x0, y0, b1, x1, y1;
    auto load = [&](int c) {
        const int k0 = c * kChunk;
        if (has0) {
            b0 = B[bo0 + k0];
            x0 = A0[ao0 + k0 * NP];
            y0 = A1[ao0 + k0 * NP];
        }
        if (has1) {
            b1 = B[bo1 + k0];
            x1 = A0[ao1 + k0 * NP];
            y1 = A1[ao1 + k0 * NP];
        }
    };
    auto stash = [&](int buf) {
        if (has0) {
            sm.bs[buf][e0 % kChunk][e0 / kChunk] = b0;
            sm.as0[buf][e0 / NP][e0 % NP] = x0;
            sm.as1[buf][e0 / NP][e0 % NP] = y0;
        }
        if (has1) {
            sm.bs[buf][e1 % kChunk][e1 / kChunk] = b1;
            sm.as0[buf][e1 / NP][e1 % NP] = x1;
            sm.as1[buf][e1 / NP][e1 % NP] = y1;
        }
    };

    load(0);
    stash(0);
    __syncthreads();
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        if (c + 1 < nchunks) load(c + 1);
#pragma unroll
        for (int s = 0; s < TPW; ++s) {
            const int t = wave + kWaves * s;
            if (t < TILES) {
                const int which = t / (Q * Q);
                const int tt = t - which * Q * Q;
                const int ti = tt % Q, tj = tt / Q;
                const double2(*as)[NP] = which ? sm.as1[buf] : sm.as0[buf];
#pragma unroll
                for (int ks = 0; ks < kChunk / 4; ++ks) {
                    const double2 a = as[4 * ks + lq][16 * ti + lr];
                    const double2 b = sm.bs[buf][4 * ks + lq][16 * tj + lr];
                    accR[s] = mfma16x16x4(a.x, b.x, accR[s]);
                    accR[s] = mfma16x16x4(-a.y, b.y, accR[s]);
                    accI[s] = mfma16x16x4(a.x, b.y, accI[s]);
                    accI[s] = mfma16x16x4(a.y, b.x, accI[s]);
                }
            }
        }
        if (c + 1 < nchunks) stash(buf ^ 1);
        __syncthreads();
    }

    double asum = 0.0;
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
        const int t = wave + kWaves * s;
        if (t < TILES) {
            const int which = t / (Q * Q);
            const int tt = t - which * Q * Q;
            const int ti = tt % Q, tj = tt / Q;
            const int col = 16 * tj + lr;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 16 * ti + 4 * v + lq;
                const int idx = row + col * NP;
                const double2 val = make_double2(accR[s][v], accI[s][v]);
                if (MODE == 0) {
                    (which ? O1 : O0)[idx] = val;
                } else if (which == 0) {
                    E1[idx] = cadd(E1[idx], val);
                    E2[idx] = cadd(E2[idx], val);
                } else {
                    O1[idx] = val;
                    if (MODE == 2) asum += hypot(val.x, val.y);
                }
            }
        }
    }
    if (MODE == 2) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) asum += __shfl_xor(asum, off, 64);
        if (lane == 0) sm.red[wave] = asum;
        __syncthreads();
        double tot = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) tot += sm.red[w];
        return tot;
    }
    return 0.0;
}

template <int Q>
__global__ void __launch_bounds__(kBlock, 1) decimate_f64_kernel(DecimateArgs a) {
    constexpr int NP = 16 * Q;
    constexpr int NN = NP * NP;
    __shared__ DecimSmem<Q> sm;
    const int tid = threadIdx.x;
    double2 *w = a.work + (size_t)blockIdx.x * kMats * NN;
    const int n = a.n;

    for (;;) {
        if (uniform(tid >> 6) == 0) {
            // wave 0 takes one ticket: lane 0 adds 1, the others add 0
            const unsigned old = atomicAdd(a.queue, (tid & 63) == 0 ? 1u : 0u);
            sm.inst = uniform((int)old);
        }
        __syncthreads();
        const int t = uniform(sm.inst);
        __syncthreads();
        if (t >= a.ninst) break;

        int ik;
        double om;
        if (a.inst_k) {
            ik = a.inst_k[t];
            om = a.inst_w[t];
        } else {
            ik = t / a.nomega;
            om = a.omega[t - ik * a.nomega];
        }
        const double eta = a.eta ? a.eta[t] : a.eta0;
        const double2 z = make_double2(om, eta);
        const double2 *H00 = a.hk + (size_t)ik * 2 * NN;
        const double2 *H01 = H00 + NN;

        double2 *Acur = w + M_A0 * NN, *Anext = w + M_A1 * NN;
        double2 *Bcur = w + M_B0 * NN, *Bnext = w + M_B1 * NN;
        double2 *EI = w + M_EI * NN, *ES = w + M_ES * NN, *ET = w + M_ET * NN;
        double2 *M1 = w + M_M1 * NN, *M2 = w + M_M2 * NN, *G = w + M_G * NN;

        // eps_i = eps_s = eps_t = H00, alpha = H01, beta = H01^dagger
        for (int e = tid; e < NN; e += kBlock) {
            const double2 h = H00[e];
            EI[e] = h;
            ES[e] = h;
            ET[e] = h;
            Acur[e] = H01[e];
            const int i = e % NP, j = e / NP;
            const double2 ht = H01[j + i * NP];
            Bcur[e] = make_double2(ht.x, -ht.y);
        }
        __syncthreads();

        int status = 0, it;
        for (it = 1; it <= a.itermax; ++it) {
            if (gj_inverse<Q>(sm, EI, z, n, G, nullptr, 0)) status |= 1;
            dual_gemm<Q, 0>(sm, Acur, Bcur, G, n, M1, M2, nullptr, nullptr);
            __syncthreads();
            dual_gemm<Q, 1>(sm, M1, M2, Bcur, n, nullptr, Bnext, EI, ES);
            __syncthreads();
            const double s = dual_gemm<Q, 2>(sm, M2, M1, Acur, n, nullptr, Anext, EI, ET);
            __syncthreads();
            double2 *tmp = Acur;
            Acur = Anext;
            Anext = tmp;
            tmp = Bcur;
            Bcur = Bnext;
            Bnext = tmp;
            if (uniform(s <= a.accuracy ? 1 : 0)) break;
        }
        if (it > a.itermax) {
            it = a.itermax;
            status |= 2;
        }

        // GLL = (z - eps_s)^-1, GRR = (z - eps_t)^-1, GB = (z - eps_i)^-1
        double dl = 0.0, dr = 0.0, db = 0.0;
#pragma unroll 1
        for (int q = 0; q < 3; ++q) {
            double2 *full = a.gfull ? a.gfull + ((size_t)t * 3 + q) * n * n : nullptr;
            const double2 *eps = q == 0 ? ES : (q == 1 ? ET : EI);
            if (gj_inverse<Q>(sm, eps, z, n, nullptr, full, full ? 2 : 1)) status |= 1;
            if (tid == 0) {
                if (q == 0) {
                    for (int i = 0; i < a.ntop; ++i) dl = dl - sm.diag[a.top[i] - 1].y / a.divisor;
                } else if (q == 1) {
                    for (int i = 0; i < a.nbot; ++i)
                        dr = dr - sm.diag[n - a.nw + a.bot[i] - 1].y / a.divisor;
                } else {
                    for (int i = 0; i < n; ++i) db = db - sm.diag[i].y / a.divisor;
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            int o = t;
            if (!a.inst_k) {
                const int iw = t - ik * a.nomega;
                o = ik + a.nk * iw;
            }
            a.dos_l[o] = dl;
            a.dos_r[o] = dr;
            a.dos_bulk[o] = db;
            if (a.iters) a.iters[o] = it;
            if (a.status) a.status[o] = status;
        }
    }
}

constexpr int kMaxQ = 5;

int padded_q(int n) { return (n + 15) / 16; }
int decimate_max_q() { return kMaxQ; }

template <int Q>
static int launch_q(const DecimateArgs &a, int slots, hipStream_t s) {
    hipLaunchKernelGGL(decimate_f64_kernel<Q>, dim3(slots), dim3(kBlock), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int Q>
static int occupancy_q() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, decimate_f64_kernel<Q>, kBlock, 0) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

int decimate_f64_slots(int q, int num_cu) {
    int per = 1;
    switch (q) {
        case 1: per = occupancy_q<1>(); break;
        case 2: per = occupancy_q<2>(); break;
        case 3: per = occupancy_q<3>(); break;
        case 4: per = occupancy_q<4>(); break;
        case 5: per = occupancy_q<5>(); break;
        default: return -1;
    }
    return per * num_cu;
}

int decimate_f64_launch(const DecimateArgs &a, int q, int slots, hipStream_t s) {
    switch (q) {
        case 1: return launch_q<1>(a, slots, s);
        case 2: return launch_q<2>(a, slots, s);
        case 3: return launch_q<3>(a, slots, s);
        case 4: return launch_q<4>(a, slots, s);
        case 5: return launch_q<5>(a, slots, s);
        default: return -1;
    }
}

}  // namespace wtsurf
