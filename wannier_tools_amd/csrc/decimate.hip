// Batched Lopez-Sancho decimation on gfx950, FP64 and FP32-with-FP64-accumulation.
//
// Restates surfgreen_1985 (src/surfgreen.f90:74-134) for many (k, omega)
// instances in one persistent launch.  One workgroup (8 waves) owns one
// instance at a time and pulls the next instance id from a device counter
// when its decimation converges, so uneven iteration counts do not idle the
// chip.  Per iteration:
//
//   g = (z - eps_i)^-1                 blocked Gauss-Jordan in LDS (MFMA updates)
//   [m1 ; m2]     = [alpha ; beta] g   dual GEMM, g read straight from LDS
//   [m1 b ; m2 b] -> eps_i, eps_s += m1 b ; beta'  = m2 b    (b = beta)
//   [m2 a ; m1 a] -> eps_i, eps_t += m2 a ; alpha' = m1 a    (a = alpha)
//   stop when sum_ij |alpha'_ij| <= accuracy
//
// the reference's 1 inverse + 6 GEMMs, grouped into three products that
// each stage one shared operand through LDS.  R is the working precision of
// the inverse and the GEMMs: double reproduces the reference; float runs
// them on the FP32 MFMA while the eps accumulators, the convergence sum, the
// three final inverses and the traces stay in FP64.
#include "wtsurf_device.h"
#include "wtsurf_internal.h"

namespace wtsurf {

template <typename T>
constexpr T cmax(T a, T b) { return a > b ? a : b; }

// LDS layout of one workgroup, in bytes.  Regions reuse the same memory in
// successive phases:
//   Gauss-Jordan:       [ S = matrix, NP x LD | T = panel pivot rows ]
//   GEMM with g in LDS: [ S = g                | A0 | A1 staging     ]
//   GEMM, B staged:     [ B staging ...        | A0 | A1 staging     ]
//   final FP64 inverse: [ S64                  | T64                 ]
template <int Q, typename R>
struct Lay {
    using C = typename Prec<R>::C;
    static constexpr int NP = 16 * Q, LD = NP + 1;
    static constexpr size_t S = (size_t)NP * LD * sizeof(C);
    static constexpr size_t T = (size_t)kPanel * NP * sizeof(C);
    static constexpr size_t A = (size_t)2 * kChunk * NP * sizeof(C);
    static constexpr size_t B = (size_t)2 * kChunk * (NP + 1) * sizeof(C);
    static_assert(B <= S, "B staging must fit under the matrix region");
    static constexpr size_t fin = (size_t)NP * LD * 16 + (size_t)kPanel * NP * 16;
    static constexpr size_t bytes = cmax(cmax(S + T, S + 2 * A), fin);
};

template <int Q>
struct Common {
    double2 diag[16 * Q];
    int pivrow[16 * Q];   // pivot row of elimination step k
    int pivstep[16 * Q];  // elimination step at which row i was the pivot (-1: not yet)
    double red[kWaves];
    int sing;
    int inst;
};

// ---------------------------------------------------------------------------
// Blocked Gauss-Jordan inverse of X = z*1 - E (the first n x n block of the
// padded FP64 matrix E), in precision R, with partial pivoting on |re|+|im|
// (the zgetrf rule; ties go to the lowest row).  X lives in LDS as
// S[col][row].  Columns are processed in panels of kPanel: wave 0 eliminates
// the panel columns in registers (one or two rows per lane, no workgroup
// barrier per step; the pivot search is a DPP wave maximum plus a ballot) and
// records the panel's pivot rows of the other columns in T; then every wave
// applies the panel's row operations to the other columns as one
// rank-kPanel MFMA update
//
//   S[:, ~C] = S0[:, ~C] + S[:, C] T      (rows pivoted in this panel: = S[:, C] T)
//
// Rows are never swapped: step k's pivot is row pivrow[k], and at the end
// X^-1[i][j] = S[pivstep[j]][pivrow[i]].
//
// out_mode 1: the diagonal of X^-1 into cm.diag (FP64).
// out_mode 2: the diagonal and the unpadded n x n X^-1 into `full` (FP64).
// out_mode 3: X^-1 in place in LDS, S[j][i] = X^-1[i][j], pad zero.
// Returns true if an exactly zero pivot was met; the result is then the
// identity, as the reference's inv hands back (zgesv skips the solve on a
// zero pivot and inv copies back the identity right-hand side,
// src/inverse.f90:44-55).
template <int Q, typename R>
__device__ bool gj_inverse(unsigned char *lds, Common<Q> &cm, const double2 *__restrict__ E,
                           double2 z, int n, int out_mode, double2 *__restrict__ full) {
    using P = Prec<R>;
    using C = typename P::C;
    constexpr int NP = 16 * Q;
    constexpr int LD = NP + 1;
    static_assert(NP <= 128, "the panel wave holds at most two rows per lane");
    constexpr int TILES = Q * Q;
    constexpr int TPW = (TILES + kWaves - 1) / kWaves;
    const int tid = opaque_tid();
    const int wave = uniform(tid >> 6), lane = tid & 63;
    const int lr = lane & 15, lq = lane >> 4;
    C(*S)[LD] = reinterpret_cast<C(*)[LD]>(lds);
    C(*T)[NP] = reinterpret_cast<C(*)[NP]>(lds + Lay<Q, R>::S);

    for (int e = tid; e < NP * NP; e += kBlock) {
        const int i = e % NP, j = e / NP;
        C v;
        v.x = 0;
        v.y = 0;
        if (i < n && j < n) {
            const double2 x = E[i + j * NP];
            v.x = (R)((i == j ? z.x : 0.0) - x.x);
            v.y = (R)((i == j ? z.y : 0.0) - x.y);
        }
        S[j][i] = v;
    }
    for (int i = tid; i < NP; i += kBlock) cm.pivstep[i] = -1;
    if (tid == 0) cm.sing = 0;
    __syncthreads();

    const int npanel = (n + kPanel - 1) / kPanel;
    for (int pb = 0; pb < npanel; ++pb) {
        const int c0 = pb * kPanel;
        // ---- panel elimination: wave 0, rows r0 = lane and r1 = lane + 64
        if (wave == 0) {
            const int r0 = lane < NP ? lane : 0, r1 = lane + 64;
            const bool has0 = lane < n;
            const bool two = NP > 64 && uniform(n > 64 ? 1 : 0);  // second rows exist
            const bool has1 = two && r1 < n;
            C a0[kPanel], a1[kPanel];
#pragma unroll
            for (int j = 0; j < kPanel; ++j) {
                a0[j] = S[c0 + j][r0];
                if (!has0) a0[j].x = a0[j].y = 0;
                a1[j].x = a1[j].y = 0;
                if (has1) a1[j] = S[c0 + j][r1];
            }
            bool free0 = has0 && cm.pivstep[r0] < 0;
            bool free1 = has1 && cm.pivstep[has1 ? r1 : 0] < 0;
            bool sing = false;
            int jdone = 0;
#pragma unroll
            for (int j = 0; j < kPanel; ++j) {
                const int c = c0 + j;
                if (c >= n) break;
                const double v0 = free0 ? (double)(fabs(a0[j].x) + fabs(a0[j].y)) : -1.0;
                const double v1 = free1 ? (double)(fabs(a1[j].x) + fabs(a1[j].y)) : -1.0;
                const double best = wave_max_f64(two ? fmax(v0, v1) : v0);
                if (!(best > 0.0)) {  // zero (or NaN) pivot column: singular
                    sing = true;
                    break;
                }
                // lowest row attaining the maximum: rows 0..63 before 64..127
                const unsigned long long m0 = __ballot(v0 == best);
                const unsigned long long m1 = two ? __ballot(v1 == best) : 0ull;
                const int slot = uniform(m0 ? 0 : 1);
                const int pl = uniform(__ffsll((long long)(m0 ? m0 : m1)) - 1);
                const int p = pl + 64 * slot;
                // this pivot row's entries outside the panel, for the MFMA update
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int col = lane + 64 * h;
                    if (col < NP) {
                        C tv;
                        tv.x = 0;
                        tv.y = 0;
                        if (col < c0 || col >= c0 + kPanel) tv = S[col][p];
                        T[j][col] = tv;
                    }
                }
                // pivot row from lane pl, normalised: entry j -> 1/pivot, others scaled
                C rowv[kPanel];
                if (slot == 0) {
#pragma unroll
                    for (int jj = 0; jj < kPanel; ++jj) {
                        rowv[jj].x = P::readlane(a0[jj].x, pl);
                        rowv[jj].y = P::readlane(a0[jj].y, pl);
                    }
                } else {
#pragma unroll
                    for (int jj = 0; jj < kPanel; ++jj) {
                        rowv[jj].x = P::readlane(a1[jj].x, pl);
                        rowv[jj].y = P::readlane(a1[jj].y, pl);
                    }
                }
                const C r = crecip_uniform(rowv[j]);
#pragma unroll
                for (int jj = 0; jj < kPanel; ++jj) rowv[jj] = (jj == j) ? r : cmul(rowv[jj], r);
                // eliminate column c from every row (pivot row overwritten below)
                {
                    const C f = a0[j];
                    a0[j].x = a0[j].y = 0;
#pragma unroll
                    for (int jj = 0; jj < kPanel; ++jj) {
                        a0[jj].x -= f.x * rowv[jj].x - f.y * rowv[jj].y;
                        a0[jj].y -= f.x * rowv[jj].y + f.y * rowv[jj].x;
                    }
                }
                if (two) {
                    const C f = a1[j];
                    a1[j].x = a1[j].y = 0;
#pragma unroll
                    for (int jj = 0; jj < kPanel; ++jj) {
                        a1[jj].x -= f.x * rowv[jj].x - f.y * rowv[jj].y;
                        a1[jj].y -= f.x * rowv[jj].y + f.y * rowv[jj].x;
                    }
                }
                if (lane == pl) {
                    if (slot == 0) {
#pragma unroll
                        for (int jj = 0; jj < kPanel; ++jj) a0[jj] = rowv[jj];
                        free0 = false;
                    } else {
#pragma unroll
                        for (int jj = 0; jj < kPanel; ++jj) a1[jj] = rowv[jj];
                        free1 = false;
                    }
                }
                if (lane == 0) {
                    cm.pivrow[c] = p;
                    cm.pivstep[p] = c;
                }
                jdone = j + 1;
            }
            // rows of T past the last column of the matrix stay zero
            for (int j = jdone; j < kPanel; ++j)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int col = lane + 64 * h;
                    if (col < NP) {
                        C tv;
                        tv.x = 0;
                        tv.y = 0;
                        T[j][col] = tv;
                    }
                }
#pragma unroll
            for (int j = 0; j < kPanel; ++j) {
                if (lane < NP) S[c0 + j][r0] = a0[j];
                if (has1) S[c0 + j][r1] = a1[j];
            }
            if (lane == 0 && sing) cm.sing = 1;
        }
        __syncthreads();
        if (uniform(cm.sing)) break;
        // ---- rank-kPanel update of every column outside the panel (MFMA)
        const int cend = c0 + kPanel;
#pragma unroll
        for (int s = 0; s < TPW; ++s) {
            const int t = wave + kWaves * s;
            if (t < TILES) {
                const int ti = t % Q, tj = t / Q;
                typename P::V4 accR = {0, 0, 0, 0}, accI = {0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < kPanel / 4; ++ks) {
                    const C a = S[c0 + 4 * ks + lq][16 * ti + lr];
                    const C b = T[4 * ks + lq][16 * tj + lr];
                    accR = P::mfma(a.x, b.x, accR);
                    accR = P::mfma(-a.y, b.y, accR);
                    accI = P::mfma(a.x, b.y, accI);
                    accI = P::mfma(a.y, b.x, accI);
                }
                const int col = 16 * tj + lr;
                if (col < c0 || col >= cend) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = 16 * ti + P::out_row(lq, v);
                        const int ps = cm.pivstep[row];
                        const bool pivot_here = ps >= c0 && ps < cend;
                        C o = S[col][row];
                        if (pivot_here) o.x = o.y = 0;
                        o.x += accR[v];
                        o.y += accI[v];
                        S[col][row] = o;
                    }
                }
            }
        }
        __syncthreads();
    }

    const bool singular = uniform(cm.sing) != 0;
    if (out_mode == 3) {
        constexpr int RPT = (NP * NP + kBlock - 1) / kBlock;
        C hold[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int e = tid + r * kBlock;
            const int i = e % NP, j = e / NP;
            C v;
            v.x = 0;
            v.y = 0;
            if (e < NP * NP && i < n && j < n) {
                if (singular)
                    v.x = (i == j) ? 1 : 0;
                else
                    v = S[cm.pivstep[j]][cm.pivrow[i]];
            }
            hold[r] = v;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int e = tid + r * kBlock;
            if (e < NP * NP) S[e / NP][e % NP] = hold[r];
        }
    } else {
        for (int i = tid; i < n; i += kBlock) {
            const C v = singular ? C{1, 0} : S[cm.pivstep[i]][cm.pivrow[i]];
            cm.diag[i] = make_double2(v.x, v.y);
        }
        if (out_mode == 2)
            for (int e = tid; e < n * n; e += kBlock) {
                const int i = e % n, j = e / n;
                C v;
                if (singular) {
                    v.x = (i == j) ? 1 : 0;
                    v.y = 0;
                } else {
                    v = S[cm.pivstep[j]][cm.pivrow[i]];
                }
                full[i + j * n] = make_double2(v.x, v.y);
            }
    }
    __syncthreads();
    return singular;
}

// ---------------------------------------------------------------------------
// Dual GEMM sharing the right operand:  P0 = A0 * B,  P1 = A1 * B  (padded
// NP x NP in precision R, K = kdim rounded up to the chunk).  The A column
// panels (and B, unless B_LDS) are staged through LDS in chunks of kChunk
// columns, double-buffered with a register prefetch of the next chunk; with
// B_LDS the right operand is the inverse Gauss-Jordan left in LDS
// (B[k][j] = S[j][k]).  Tile t in [0, 2*Q*Q) belongs to wave t % 8;
// accumulators stay in registers for the whole K loop.
//
//   MODE 0: O0 = P0, O1 = P1
//   MODE 1: E1 += P0, E2 += P0 (E1 first, as the reference adds; FP64), O1 = P1
//   MODE 2: as MODE 1, and returns sum_ij |P1_ij| (FP64) to every thread
template <int Q, typename R, int MODE, bool B_LDS>
__device__ double dual_gemm(unsigned char *lds, Common<Q> &cm,
                            const typename Prec<R>::C *__restrict__ A0,
                            const typename Prec<R>::C *__restrict__ A1,
                            const typename Prec<R>::C *__restrict__ B, int kdim,
                            typename Prec<R>::C *__restrict__ O0, typename Prec<R>::C *__restrict__ O1,
                            double2 *__restrict__ E1, double2 *__restrict__ E2) {
    using P = Prec<R>;
    using C = typename P::C;
    constexpr int NP = 16 * Q;
    constexpr int LD = NP + 1;
    constexpr int TILES = 2 * Q * Q;
    constexpr int TPW = (TILES + kWaves - 1) / kWaves;
    const int tid = opaque_tid();
    const int wave = uniform(tid >> 6), lane = tid & 63;
    const int lr = lane & 15, lq = lane >> 4;
    C(*Sg)[LD] = reinterpret_cast<C(*)[LD]>(lds);
    C(*as0)[kChunk][NP] = reinterpret_cast<C(*)[kChunk][NP]>(lds + Lay<Q, R>::S);
    C(*as1)[kChunk][NP] = reinterpret_cast<C(*)[kChunk][NP]>(lds + Lay<Q, R>::S + Lay<Q, R>::A);
    C(*bs)[kChunk][NP + 1] = reinterpret_cast<C(*)[kChunk][NP + 1]>(lds);

    typename P::V4 accR[TPW], accI[TPW];
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
        accR[s] = typename P::V4{0, 0, 0, 0};
        accI[s] = typename P::V4{0, 0, 0, 0};
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
    C b0, x0, y0, b1, x1, y1;
    auto load = [&](int c) {
        const int k0 = c * kChunk;
        if (has0) {
            if (!B_LDS) b0 = B[bo0 + k0];
            x0 = A0[ao0 + k0 * NP];
            y0 = A1[ao0 + k0 * NP];
        }
        if (has1) {
            if (!B_LDS) b1 = B[bo1 + k0];
            x1 = A0[ao1 + k0 * NP];
            y1 = A1[ao1 + k0 * NP];
        }
    };
    auto stash = [&](int buf) {
        if (has0) {
            if (!B_LDS) bs[buf][e0 % kChunk][e0 / kChunk] = b0;
            as0[buf][e0 / NP][e0 % NP] = x0;
            as1[buf][e0 / NP][e0 % NP] = y0;
        }
        if (has1) {
            if (!B_LDS) bs[buf][e1 % kChunk][e1 / kChunk] = b1;
            as0[buf][e1 / NP][e1 % NP] = x1;
            as1[buf][e1 / NP][e1 % NP] = y1;
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
                const C(*as)[NP] = which ? as1[buf] : as0[buf];
#pragma unroll
                for (int ks = 0; ks < kChunk / 4; ++ks) {
                    const C a = as[4 * ks + lq][16 * ti + lr];
                    const C b = B_LDS ? Sg[16 * tj + lr][c * kChunk + 4 * ks + lq]
                                      : bs[buf][4 * ks + lq][16 * tj + lr];
                    accR[s] = P::mfma(a.x, b.x, accR[s]);
                    accR[s] = P::mfma(-a.y, b.y, accR[s]);
                    accI[s] = P::mfma(a.x, b.y, accI[s]);
                    accI[s] = P::mfma(a.y, b.x, accI[s]);
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
                const int row = 16 * ti + P::out_row(lq, v);
                const int idx = row + col * NP;
                C val;
                val.x = accR[s][v];
                val.y = accI[s][v];
                if (MODE == 0) {
                    (which ? O1 : O0)[idx] = val;
                } else if (which == 0) {
                    const double2 d = make_double2(val.x, val.y);
                    E1[idx] = cadd(E1[idx], d);
                    E2[idx] = cadd(E2[idx], d);
                } else {
                    O1[idx] = val;
                    if (MODE == 2) asum += hypot((double)val.x, (double)val.y);
                }
            }
        }
    }
    if (MODE == 2) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) asum += __shfl_xor(asum, off, 64);
        if (lane == 0) cm.red[wave] = asum;
        __syncthreads();
        double tot = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) tot += cm.red[w];
        return tot;
    }
    return 0.0;
}

// Per-slot workspace: alpha, alpha', beta, beta', m1, m2 in precision R,
// then eps_i, eps_s, eps_t in FP64.
template <int Q, typename R>
constexpr size_t slot_bytes() {
    constexpr size_t NN = (size_t)(16 * Q) * (16 * Q);
    return 6 * NN * sizeof(typename Prec<R>::C) + 3 * NN * sizeof(double2);
}

template <int Q, typename R>
__global__ void __launch_bounds__(kBlock, 1) decimate_kernel(DecimateArgs a) {
    using C = typename Prec<R>::C;
    constexpr int NP = 16 * Q;
    constexpr int NN = NP * NP;
    __shared__ __align__(16) unsigned char lds[Lay<Q, R>::bytes];
    __shared__ Common<Q> cm;
    const int tid = threadIdx.x;
    unsigned char *slot = reinterpret_cast<unsigned char *>(a.work) + (size_t)blockIdx.x * slot_bytes<Q, R>();
    C *wc = reinterpret_cast<C *>(slot);
    double2 *wd = reinterpret_cast<double2 *>(slot + 6 * (size_t)NN * sizeof(C));
    const int n = a.n;

    for (;;) {
        if (uniform(tid >> 6) == 0) {
            // wave 0 takes one ticket: lane 0 adds 1, the others add 0
            const unsigned old = atomicAdd(a.queue, (tid & 63) == 0 ? 1u : 0u);
            cm.inst = uniform((int)old);
        }
        __syncthreads();
        const int t = uniform(cm.inst);
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

        C *Acur = wc, *Anext = wc + NN, *Bcur = wc + 2 * NN, *Bnext = wc + 3 * NN;
        C *M1 = wc + 4 * NN, *M2 = wc + 5 * NN;
        double2 *EI = wd, *ES = wd + NN, *ET = wd + 2 * NN;

        // eps_i = eps_s = eps_t = H00, alpha = H01, beta = H01^dagger
        for (int e = tid; e < NN; e += kBlock) {
            const double2 h = H00[e];
            EI[e] = h;
            ES[e] = h;
            ET[e] = h;
            const double2 h1 = H01[e];
            C av;
            av.x = (R)h1.x;
            av.y = (R)h1.y;
            Acur[e] = av;
            const int i = e % NP, j = e / NP;
            const double2 ht = H01[j + i * NP];
            C bv;
            bv.x = (R)ht.x;
            bv.y = (R)(-ht.y);
            Bcur[e] = bv;
        }
        __syncthreads();

        int status = 0, it;
        for (it = 1; it <= a.itermax; ++it) {
            if (gj_inverse<Q, R>(lds, cm, EI, z, n, 3, nullptr)) status |= 1;
            dual_gemm<Q, R, 0, true>(lds, cm, Acur, Bcur, nullptr, n, M1, M2, nullptr, nullptr);
            __syncthreads();
            dual_gemm<Q, R, 1, false>(lds, cm, M1, M2, Bcur, n, nullptr, Bnext, EI, ES);
            __syncthreads();
            const double s = dual_gemm<Q, R, 2, false>(lds, cm, M2, M1, Acur, n, nullptr, Anext, EI, ET);
            __syncthreads();
            C *tmp = Acur;
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

        // GLL = (z - eps_s)^-1, GRR = (z - eps_t)^-1, GB = (z - eps_i)^-1, in FP64
        double dl = 0.0, dr = 0.0, db = 0.0;
#pragma unroll 1
        for (int q = 0; q < 3; ++q) {
            double2 *full = a.gfull ? a.gfull + ((size_t)t * 3 + q) * n * n : nullptr;
            const double2 *eps = q == 0 ? ES : (q == 1 ? ET : EI);
            if (gj_inverse<Q, double>(lds, cm, eps, z, n, full ? 2 : 1, full)) status |= 1;
            if (tid == 0) {
                if (q == 0) {
                    for (int i = 0; i < a.ntop; ++i) dl = dl - cm.diag[a.top[i] - 1].y / a.divisor;
                } else if (q == 1) {
                    for (int i = 0; i < a.nbot; ++i)
                        dr = dr - cm.diag[n - a.nw + a.bot[i] - 1].y / a.divisor;
                } else {
                    for (int i = 0; i < n; ++i) db = db - cm.diag[i].y / a.divisor;
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

template <int Q, typename R>
static int launch_q(const DecimateArgs &a, int slots, hipStream_t s) {
    hipLaunchKernelGGL((decimate_kernel<Q, R>), dim3(slots), dim3(kBlock), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int Q, typename R>
static int occupancy_q() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, decimate_kernel<Q, R>, kBlock, 0) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <typename R>
static int dispatch_slots(int q) {
    switch (q) {
        case 1: return occupancy_q<1, R>();
        case 2: return occupancy_q<2, R>();
        case 3: return occupancy_q<3, R>();
        case 4: return occupancy_q<4, R>();
        case 5: return occupancy_q<5, R>();
        default: return -1;
    }
}

template <typename R>
static size_t dispatch_slot_bytes(int q) {
    switch (q) {
        case 1: return slot_bytes<1, R>();
        case 2: return slot_bytes<2, R>();
        case 3: return slot_bytes<3, R>();
        case 4: return slot_bytes<4, R>();
        case 5: return slot_bytes<5, R>();
        default: return 0;
    }
}

template <typename R>
static int dispatch_launch(const DecimateArgs &a, int q, int slots, hipStream_t s) {
    switch (q) {
        case 1: return launch_q<1, R>(a, slots, s);
        case 2: return launch_q<2, R>(a, slots, s);
        case 3: return launch_q<3, R>(a, slots, s);
        case 4: return launch_q<4, R>(a, slots, s);
        case 5: return launch_q<5, R>(a, slots, s);
        default: return -1;
    }
}

int decimate_slots(int q, int precision, int num_cu) {
    const int per = precision == 0 ? dispatch_slots<double>(q) : dispatch_slots<float>(q);
    return per > 0 ? per * num_cu : -1;
}

size_t decimate_slot_bytes(int q, int precision) {
    return precision == 0 ? dispatch_slot_bytes<double>(q) : dispatch_slot_bytes<float>(q);
}

int decimate_launch(const DecimateArgs &a, int q, int precision, int slots, hipStream_t s) {
    return precision == 0 ? dispatch_launch<double>(a, q, slots, s)
                          : dispatch_launch<float>(a, q, slots, s);
}

}  // namespace wtsurf
