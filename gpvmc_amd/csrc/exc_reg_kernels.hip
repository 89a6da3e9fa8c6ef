// gfx950 register-resident channel (exciton) walk: one 512-thread workgroup per walker.
//
// The chain, estimator and slot bookkeeping are those of the streamed channel kernel
// (LatticeStepper::trystep/step on a channel, LatticeStepper.cpp:36-147;
// SlaterDeterminant::VirtUpdate/Update, SlaterDeterminant.cpp:96-326; linalg::DetUpdate
// and InvUpdate, linalg.cpp:107-573; ProjHeis::measure, ProjHeis.cpp:40-154).  Here each
// walker's matrices stay in VGPRs for a whole launch:
//
//   M_fl = [Ai_fl ; Q_fl],   Ai_fl[r][f] = (A_fl^-1)(label r, slot f),
//                            Q_fl[q][f]  = sum_r Phi_fl(o_q, site r) Ai_fl[r][f],
//
// o_q running over the orbitals that some hop from the current state fills (the rows
// rho_m Ai of the capacitance).  Rows are cut in blocks of 16 and columns in blocks of 4,
// held in the A-operand layout of v_mfma_*_16x16x4 (lane l: row l % 16, column l / 16 of
// each 16 x 4 tile), so the ProjHeis contraction Z = M x Phi(current orbitals, sites of
// the other flavour) runs on MFMA straight from the registers.
//
// Capacitance of a move of particle r to site s with the orbital hops o_m -> o'_m in
// slots f_m (c = Phi(current orbitals, s), y = Ai c, KG = Q c):
//   B = [[ y[r],                 Ai[r][f_n] ],
//        [ KG_m - Phi(o'_m, s),  Q_m[f_n]   ]].
// This is the streamed kernel's capacitance with the multiples D_m of row 0 taken out of
// rows m: a row operation, which changes neither det B nor B^-1 V (V's rows m carry the
// same multiple of V's row 0).
//
// Accepted move: Ai' = Ai - U W with W = B^-1 V, U = [y - e_r, Ai[:, f_m]],
// V = [Ai[r, :]; Q_m - e_{f_m}].  The Q rows follow from Q'_q = Phi'(o_q) Ai' with
// Phi'(o_q) = Phi(o_q) + delta_q e_r^T, delta_q = Phi(o_q, s) - Phi(o_q, site r):
//   Q'_q = Q_q + delta_q V_0 - sum_a kappa_a W_a,
//   kappa_0 = KG_q - Phi(o_q, s) + delta_q y[r],   kappa_m = Q_q[f_m] + delta_q Ai[r][f_m].
// The row of a filled orbital o'_m is taken over by the emptied o_m, whose row before the
// move is e_{f_m}: the same formula with Q_q = e_{f_m}, KG_q - Phi(o_q, s) = 0 and
// delta = Phi(o_m, s) - Phi(o_m, site r).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "exc_common.h"
#include "vmc_kernels.h"

namespace gr {

using gv::ExcArgs;
using namespace gz;

template <typename T> struct Mf;
template <> struct Mf<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __device__ static v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    // row of output element v of lane l; its column is l % 16
    __device__ static int orow(int l, int v) { return 4 * (l >> 4) + v; }
};
template <> struct Mf<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    __device__ static v4 mma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int orow(int l, int v) { return (l >> 4) + 4 * v; }
};

// eight waves per walker, two per SIMD
constexpr int NW = 8, NT = 64 * NW;

struct Sizes {
    int Ns, n0, n1, nmax, nq0, nq1, nqm, rb0, rb1, r16, nc, nk, zt, ngrp;
};

__host__ __device__ __forceinline__ Sizes make_sizes(int L, int n0, int n1, int nq0, int nq1, int rb0, int rb1, int cb, int nk)
{
    Sizes z;
    z.Ns = L * L;
    z.n0 = n0;
    z.n1 = n1;
    z.nmax = n0 > n1 ? n0 : n1;
    z.nq0 = nq0;
    z.nq1 = nq1;
    z.nqm = nq0 > nq1 ? nq0 : nq1;
    z.rb0 = rb0;
    z.rb1 = rb1;
    z.r16 = 16 * (rb0 > rb1 ? rb0 : rb1);
    z.nc = 4 * cb;
    z.nk = nk;
    z.zt = 16 * ((z.nmax + 15) / 16);
    const int g = NT / nk;
    z.ngrp = g < 1 ? 1 : (g > 8 ? 8 : g);
    return z;
}

// LDS of one walker (byte offsets); ts = bytes of one complex element of the working
// precision
struct Off {
    size_t site, nbr, part, cur, slot, qrow, qorb, rep, c2, ra, alpha, dets, a0, hpart, binv, red, ired, hop, un, mz,
        total;
    bool qboth;  // the shared area holds both flavours' Q rows at once (one RA build for both)
};

__host__ __device__ __forceinline__ size_t take(size_t &off, size_t bytes)
{
    const size_t o = off;
    off += (bytes + 15) & ~(size_t)15;
    return o;
}

// measurement with its tiles in LDS: the Q rows' products Zq_fl[target][q] of both
// flavours, the inverse-row entries y0[fl][target][bond direction] and the packed neighbour
// map nbt[fl][target]
__host__ __device__ __forceinline__ size_t mz_bytes(const Sizes &z, size_t ts)
{
    return ts * z.zt * (size_t)(z.nq0 + z.nq1) + ts * 8 * (size_t)z.nmax + 4 * 2 * (size_t)z.nmax;
}

__host__ __device__ __forceinline__ Off carve(const Sizes &z, size_t ts, bool lz = false)
{
    Off o;
    size_t off = 0;
    o.site = take(off, 4 * (size_t)z.Ns);
    o.nbr = take(off, 16 * (size_t)z.Ns);
    o.part = take(off, 8 * (size_t)z.nmax);
    o.cur = take(off, 8 * (size_t)z.nmax);
    o.slot = take(off, 8 * (size_t)z.Ns);
    o.qrow = take(off, 8 * (size_t)z.Ns);
    o.qorb = take(off, 8 * (size_t)z.nqm);
    o.rep = take(off, 8 * (size_t)z.r16);
    // s_c[2 nc] s_row[2 nc] s_z[2 r16] colb[4 r16] vq[4 nc] wr[6 nc] kap[6 r16] dl[2 r16]
    o.c2 = take(off, ts * (14 * (size_t)z.nc + 14 * (size_t)z.r16));
    o.ra = take(off, 16 * 8 * (size_t)z.nk);
    o.alpha = take(off, 16 * (size_t)z.nk);
    o.dets = take(off, 32 * (size_t)z.nk);
    o.a0 = take(off, 16 * (size_t)z.nk);
    o.hpart = take(off, 16 * (size_t)z.nk * z.ngrp);
    o.binv = take(off, 16 * 18);
    o.red = take(off, 16 * 16);
    o.ired = take(off, 4 * 16);
    o.hop = take(off, 4 * 19 * (size_t)z.nk);
    // the shared area: the dump of the Q rows for the RA blocks (both flavours at once where
    // that still fits in 160 KB with the rest, else one flavour at a time), the measurement's
    // Phi tiles and bond list
    const size_t base = off;
    auto umax = [](size_t x, size_t y) { return x > y ? x : y; };
    for (int both = 1; both >= 0; --both) {
        off = base;
        const size_t rows = both ? (size_t)(z.nq0 + z.nq1) : (size_t)z.nqm;
        o.un = take(off, ts * umax(rows * z.nc, (size_t)z.nc * 16));
        o.mz = take(off, lz ? mz_bytes(z, ts) : 0);
        o.total = off;
        o.qboth = both != 0;
        if (!both || o.total + 64 <= 160 * 1024) break;
    }
    return o;
}

template <typename C2> __device__ __forceinline__ C2 mk2(double x, double y)
{
    C2 v;
    v.x = x;
    v.y = y;
    return v;
}

// MK = 0, mode 0: walk n_steps (no measurement: the host launches the measurements
// between walk segments); 2: initial state choice (largest |a_k|^2, vmc.cpp:243-254) and
// weight; 3: weight only; 4: amplitudes into amp_out.
// MK = 1: one ProjHeis measurement, into the sums (mode 5) or only the traces (mode 1);
// MK = 2: the same with the contraction's tiles in LDS instead of the HBM scratch zscr.
template <typename T, int CB, int RBW, int MK>
__global__ void __launch_bounds__(NT) excr_kernel(ExcArgs a)
{
    using C2 = typename cpx<T>::t;
    using MF = Mf<T>;
    constexpr int NC = 4 * CB;
    extern __shared__ __align__(16) char smem[];
    const int w = a.order ? a.order[blockIdx.x] : (int)blockIdx.x;
    // lane ids are re-declared opaque at every step (see the walk loop) so that the
    // compiler recomputes lane-dependent addresses instead of holding them in VGPRs
    int tid = threadIdx.x, lane = tid & 63, li = lane & 15, kq = lane >> 4;
    // wave-uniform values go through readfirstlane so that branches on them are scalar
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const Sizes z = make_sizes(a.L, a.n_up, a.n_dn, a.nq0, a.nq1, a.rb0, a.rb1, CB, a.nk);
    const int Ns = z.Ns, nk = z.nk, nmax = z.nmax, nqm = z.nqm, r16 = z.r16, RB0 = z.rb0;
    const int RBT = z.rb0 + z.rb1, RS = 16 * RBT;
    auto NF = [&](int fl) { return fl ? z.n1 : z.n0; };
    auto NQ = [&](int fl) { return fl ? z.nq1 : z.nq0; };
    constexpr bool LZ = MK == 2;
    const Off lo = carve(z, sizeof(C2), LZ);
    int *const s_site = reinterpret_cast<int *>(smem + lo.site);
    int *const s_part = reinterpret_cast<int *>(smem + lo.part);
    int *const s_cur = reinterpret_cast<int *>(smem + lo.cur);
    int *const s_slot = reinterpret_cast<int *>(smem + lo.slot);
    int *const s_qrow = reinterpret_cast<int *>(smem + lo.qrow);
    int *const s_qorb = reinterpret_cast<int *>(smem + lo.qorb);
    int *const s_rep = reinterpret_cast<int *>(smem + lo.rep);
    zc *const s_ra = reinterpret_cast<zc *>(smem + lo.ra);
    zc *const s_alpha = reinterpret_cast<zc *>(smem + lo.alpha);
    zc *const s_dets = reinterpret_cast<zc *>(smem + lo.dets);
    zc *const s_a0 = reinterpret_cast<zc *>(smem + lo.a0);
    zc *const s_hpart = reinterpret_cast<zc *>(smem + lo.hpart);
    zc *const s_binv = reinterpret_cast<zc *>(smem + lo.binv);
    zc *const s_red = reinterpret_cast<zc *>(smem + lo.red);
    int *const s_ired = reinterpret_cast<int *>(smem + lo.ired);
    int *const s_nbr = reinterpret_cast<int *>(smem + lo.nbr);  // the lattice's neighbour table
    char *const s_un = smem + lo.un;
    // thread 0: the product of the accepted moves' determinant ratios per flavour, mantissa
    // [0..3] and binary exponent [4..5], folded into logabs / phase when the launch ends (in
    // the spare bytes past the carved layout)
    double *const s_amp = reinterpret_cast<double *>(smem + lo.total);
    // the current state's row of the hop graph, resolved to Q rows and slots, state index
    // innermost (the lanes of a wave are consecutive states): hn[fl][k] hops,
    // hto/hfr[fl][m][k] orbitals, hq[fl][m][k] Q row of hto, hs[fl][m][k] slot of hfr,
    // hsg[k] HopSign
    int *const s_hn = reinterpret_cast<int *>(smem + lo.hop);
    int *const s_hto = s_hn + 2 * nk;
    int *const s_hfr = s_hto + 4 * nk;
    int *const s_hq = s_hfr + 4 * nk;
    int *const s_hs = s_hq + 4 * nk;
    int *const s_hsg = s_hs + 4 * nk;
    C2 *s_c = reinterpret_cast<C2 *>(smem + lo.c2);
    C2 *s_row = s_c + 2 * NC;
    C2 *s_z = s_row + 2 * NC;
    C2 *s_colb = s_z + 2 * r16;
    C2 *s_vq = s_colb + 4 * r16;
    C2 *s_wr = s_vq + 4 * NC;
    C2 *s_kap = s_wr + 6 * NC;
    C2 *s_dl = s_kap + 6 * r16;

    const size_t chb = (size_t)(a.chan ? a.chan[w] : 0) * nk;
    auto ck = [&](int c, int k) { return (chb + c) * nk + k; };
    auto hop_n = [&](int c, int k, int fl) { return a.hop_n[ck(c, k) * 2 + fl]; };
    auto hop_from = [&](int c, int k, int fl, int h) { return a.hop_orb[((ck(c, k) * 2 + fl) * 2 + h) * 2]; };
    auto hop_to = [&](int c, int k, int fl, int h) { return a.hop_orb[((ck(c, k) * 2 + fl) * 2 + h) * 2 + 1]; };
    auto hsign = [&](int c, int k) { return a.hopsign[ck(c, k)]; };
    const C2 *tabT = reinterpret_cast<const C2 *>(a.tabT) + (size_t)a.group[w] * 2 * Ns * Ns;
    auto phi = [&](int fl, int orb, int site) { return toz(tabT[((size_t)fl * Ns + site) * Ns + orb]); };
    C2 *Mg = reinterpret_cast<C2 *>(a.M) + (size_t)w * RS * NC;

    // ---- walker state into LDS and registers
    for (int i = tid; i < 4 * Ns; i += NT) s_nbr[i] = a.nbr[i];
    for (int i = tid; i < Ns; i += NT) {
        s_site[i] = a.site[(size_t)w * Ns + i];
        s_slot[i] = -1;
        s_slot[Ns + i] = -1;
        s_qrow[i] = -1;
        s_qrow[Ns + i] = -1;
    }
    for (int i = tid; i < 2 * nmax; i += NT) {
        s_part[i] = a.part[(size_t)w * 2 * nmax + i];
        s_cur[i] = a.cur[(size_t)w * 2 * nmax + i];
    }
    for (int i = tid; i < 2 * nqm; i += NT) s_qorb[i] = a.qorb[(size_t)w * 2 * nqm + i];
    // s_c, s_row (adjacent) and the W rows start at zero: the branch-free update multiplies
    // them by zero coefficients before the first move writes them
    for (int i = tid; i < 4 * NC; i += NT) s_c[i] = mk2<C2>(0.0, 0.0);
    for (int i = tid; i < 6 * NC; i += NT) s_wr[i] = mk2<C2>(0.0, 0.0);
    int cs = __builtin_amdgcn_readfirstlane(a.cstate[w]);
    __syncthreads();
    for (int i = tid; i < 2 * nmax; i += NT) {
        const int fl = i / nmax, f = i - fl * nmax;
        if (f < NF(fl)) s_slot[fl * Ns + s_cur[i]] = f;
    }
    for (int i = tid; i < 2 * nqm; i += NT) {
        const int fl = i / nqm, q = i - fl * nqm;
        if (q < NQ(fl) && s_qorb[i] >= 0) s_qrow[fl * Ns + s_qorb[i]] = q;
    }
    T mr[RBW][CB], mi[RBW][CB];
#pragma unroll
    for (int j = 0; j < RBW; ++j) {
        const int rbg = wv + j * NW;
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
            C2 v = mk2<C2>(0.0, 0.0);
            if (rbg < RBT) v = Mg[(size_t)(rbg * 16 + li) * NC + 4 * cb + kq];
            mr[j][cb] = v.x;
            mi[j][cb] = v.y;
        }
    }
    __syncthreads();

    // row block j of this wave: flavour and row within the flavour (false past the end)
    auto rinfo = [&](int j, int &fl, int &row) __attribute__((always_inline)) {
        const int rbg = wv + j * NW;
        if (rbg >= RBT) return false;
        fl = rbg >= RB0 ? 1 : 0;
        row = (rbg - (fl ? RB0 : 0)) * 16 + li;
        return true;
    };
    auto store_regs = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < RBW; ++j) {
            const int rbg = wv + j * NW;
            if (rbg < RBT) {
#pragma unroll
                for (int cb = 0; cb < CB; ++cb)
                    Mg[(size_t)(rbg * 16 + li) * NC + 4 * cb + kq] = mk2<C2>(mr[j][cb], mi[j][cb]);
            }
        }
    };
    auto cache_hops = [&]() __attribute__((always_inline)) {
        for (int e = tid; e < nk * 2; e += NT) {
            const int fl = e >= nk ? 1 : 0, k = e - fl * nk;
            const int n = hop_n(cs, k, fl);
            s_hn[e] = n;
            if (fl == 0) s_hsg[k] = hsign(cs, k);
            for (int m = 0; m < 2; ++m) {
                int o1 = -1, o2 = -1, q = 0, sl = 0;
                if (m < n) {
                    o1 = hop_from(cs, k, fl, m);
                    o2 = hop_to(cs, k, fl, m);
                    q = s_qrow[fl * Ns + o2];
                    sl = s_slot[fl * Ns + o1];
                }
                s_hfr[(fl * 2 + m) * nk + k] = o1;
                s_hto[(fl * 2 + m) * nk + k] = o2;
                s_hq[(fl * 2 + m) * nk + k] = q;
                s_hs[(fl * 2 + m) * nk + k] = sl;
            }
        }
        __syncthreads();
    };
    // RA[fl][m][n][k] = Q_fl[row of hop_to(c, k, fl, m)][slot of hop_from(c, k, fl, n)]
    auto build_ra = [&]() __attribute__((always_inline)) {
        C2 *Qd = reinterpret_cast<C2 *>(s_un);
        if (lo.qboth) {
            // both flavours' Q rows (flavour 1 after flavour 0's nq0 rows), one barrier, then
            // every RA entry of both flavours
#pragma unroll
            for (int j = 0; j < RBW; ++j) {
                int fl, row;
                if (rinfo(j, fl, row) && row >= NF(fl) && row < NF(fl) + NQ(fl)) {
                    C2 *qd = Qd + (fl ? (size_t)z.nq0 * NC : 0);
#pragma unroll
                    for (int cb = 0; cb < CB; ++cb)
                        qd[(row - NF(fl)) * NC + 4 * cb + kq] = mk2<C2>(mr[j][cb], mi[j][cb]);
                }
            }
            __syncthreads();
            for (int e = tid; e < nk * 8; e += NT) {
                const int fmq = e / nk, k = e - fmq * nk, fl = fmq >> 2, m = (fmq >> 1) & 1, q = fmq & 1;
                const int R = s_hn[fl * nk + k];
                const C2 *qd = Qd + (fl ? (size_t)z.nq0 * NC : 0);
                zc v = zmk(0.0, 0.0);
                if (m < R && q < R) v = toz(qd[s_hq[(fl * 2 + m) * nk + k] * NC + s_hs[(fl * 2 + q) * nk + k]]);
                s_ra[((fl * 2 + m) * 2 + q) * nk + k] = v;
            }
            __syncthreads();
            return;
        }
        for (int fl = 0; fl < 2; ++fl) {
            if (NQ(fl) > 0) {
#pragma unroll
                for (int j = 0; j < RBW; ++j) {
                    int f_, row;
                    if (rinfo(j, f_, row) && f_ == fl && row >= NF(fl) && row < NF(fl) + NQ(fl)) {
#pragma unroll
                        for (int cb = 0; cb < CB; ++cb)
                            Qd[(row - NF(fl)) * NC + 4 * cb + kq] = mk2<C2>(mr[j][cb], mi[j][cb]);
                    }
                }
            }
            __syncthreads();
            for (int e = tid; e < nk * 4; e += NT) {
                const int mq = e / nk, k = e - mq * nk, m = mq >> 1, q = mq & 1;
                const int R = s_hn[fl * nk + k];
                zc v = zmk(0.0, 0.0);
                if (m < R && q < R)
                    v = toz(Qd[s_hq[(fl * 2 + m) * nk + k] * NC + s_hs[(fl * 2 + q) * nk + k]]);
                s_ra[((fl * 2 + m) * 2 + q) * nk + k] = v;
            }
            __syncthreads();
        }
    };
    // sum_k |alpha_k|^2 and the first k of largest |alpha_k|^2 (std::max_element) from each
    // thread's partials: DPP wave reductions into lane 63, then one barrier and a pass over
    // the waves' results.  The results alternate between two halves of s_red, so the next
    // reduction writes the other half; the one after it comes after this one's barrier in
    // every thread.
    int rpar = 0;
    auto reduce_partials = [&](double ls, double bv, int bk, double &sum, int &kmax) __attribute__((always_inline)) {
        wave_sum_argmax_to_lane63(ls, bv, bk);
        zc *const rd = s_red + rpar * NW;
        int *const ri = s_ired + rpar * NW;
        rpar ^= 1;
        if (lane == 63) {
            rd[wv] = zmk(ls, bv);
            ri[wv] = bk;
        }
        __syncthreads();
        sum = 0.0;
        bv = -1.0;
        bk = nk;
        for (int q = 0; q < NW; ++q) {
            sum += rd[q].x;
            if (rd[q].y > bv || (rd[q].y == bv && ri[q] < bk)) {
                bv = rd[q].y;
                bk = ri[q];
            }
        }
        kmax = __builtin_amdgcn_readfirstlane(bk);
    };
    auto reduce_alpha = [&](double &sum, int &kmax) __attribute__((always_inline)) {
        double ls = 0.0, bv = -1.0;
        int bk = nk;
        for (int k = tid; k < nk; k += NT) {
            const double v = znorm(s_alpha[k]);
            ls += v;
            if (v > bv) {
                bv = v;
                bk = k;
            }
        }
        reduce_partials(ls, bv, bk, sum, kmax);
    };
    // amplitudes of every state at the current configuration, relative to the current state
    auto eval0 = [&]() __attribute__((always_inline)) {
        for (int k = tid; k < nk; k += NT) {
            zc tot = zmk((double)s_hsg[k], 0.0);
#pragma unroll
            for (int fl = 0; fl < 2; ++fl) {
                const int R = s_hn[fl * nk + k];
                zc B[9];
                eye3(B);
#pragma unroll
                for (int m = 0; m < 2; ++m)
#pragma unroll
                    for (int q = 0; q < 2; ++q)
                        if (m < R && q < R) B[m * 3 + q] = s_ra[((fl * 2 + m) * 2 + q) * nk + k];
                const zc d = det3(B);
                s_dets[k * 2 + fl] = d;
                tot = zmul(tot, d);
            }
            s_alpha[k] = tot;
        }
        __syncthreads();
    };
    // capacitance of state k for the move (ci -> sb, cj -> sa), flavour fl, from the
    // step's LDS vectors
    // (with usepf: pf0, pf1 = Phi(hop_to(m), snew) of state k's hops, loaded ahead by the caller)
    auto cap_move = [&](int k, int fl, int ci, int sa, int sb, int cj, zc *B, bool usepf = false,
                        C2 pf0 = C2(), C2 pf1 = C2()) __attribute__((always_inline)) {
        const int R = s_hn[fl * nk + k], r = fl ? cj : ci, snew = fl ? sa : sb;
        eye3(B);
        B[0] = toz(s_z[fl * r16 + r]);
#pragma unroll
        for (int q = 0; q < 2; ++q)
            if (q < R) B[1 + q] = toz(s_row[fl * NC + s_hs[(fl * 2 + q) * nk + k]]);
#pragma unroll
        for (int m = 0; m < 2; ++m)
            if (m < R) {
                const int o2 = s_hto[(fl * 2 + m) * nk + k];
                B[(1 + m) * 3] = zsub(toz(s_z[fl * r16 + NF(fl) + s_hq[(fl * 2 + m) * nk + k]]),
                                      usepf ? toz(m ? pf1 : pf0) : phi(fl, o2, snew));
#pragma unroll
                for (int q = 0; q < 2; ++q)
                    if (q < R) B[(1 + m) * 3 + 1 + q] = s_ra[((fl * 2 + m) * 2 + q) * nk + k];
            }
    };
    // ratios of every state for the swap moving up particle ci to sb and down particle
    // cj to sa: c vectors, one GEMV per flavour, the moving rows, the capacitances
    auto eval_mv = [&](int ci, int sa, int sb, int cj, double &S2, int &kn) __attribute__((always_inline)) {
        // Phi(hop_to(m), snew) of the hops of thread tid's (state, flavour) = (tid / 2, tid % 2)
        // (the first pass of the capacitance loop), loaded now so that their latency passes
        // under the c vectors and the GEMV (where the matrix leaves the registers for it: up
        // to 38 f32 or 19 f64 tiles per lane)
        constexpr bool PF = CB * RBW * (int)sizeof(T) <= 152;
        C2 pf[2];
        if constexpr (PF) {
            const int e = tid < 2 * nk ? tid : 0, kp = e >> 1, fp = e & 1;
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const int o2 = s_hto[(fp * 2 + m) * nk + kp];
                pf[m] = tabT[((size_t)fp * Ns + (fp ? sa : sb)) * Ns + (o2 >= 0 ? o2 : 0)];
            }
        }
        for (int t = tid; t < 2 * NC; t += NT) {
            const int fl = t >= NC ? 1 : 0, f = t - fl * NC;
            C2 v = mk2<C2>(0.0, 0.0);
            if (f < NF(fl)) v = tabT[((size_t)fl * Ns + (fl ? sa : sb)) * Ns + s_cur[fl * nmax + f]];
            s_c[t] = v;
        }
        __syncthreads();
        {
            // column blocks outer, row blocks inner: each c element is used at once
            T ar[RBW], ai[RBW];
            int flj[RBW], rowj[RBW];
            bool okj[RBW];
#pragma unroll
            for (int j = 0; j < RBW; ++j) {
                flj[j] = 0;
                rowj[j] = 0;
                okj[j] = rinfo(j, flj[j], rowj[j]);
                ar[j] = 0;
                ai[j] = 0;
            }
            // no branch in the column loop (rows past the end are zero in registers), so
            // the LDS loads of successive column blocks overlap
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) {
#pragma unroll
                for (int j = 0; j < RBW; ++j) {
                    const C2 c = s_c[flj[j] * NC + 4 * cb + kq];
                    ar[j] += mr[j][cb] * c.x - mi[j][cb] * c.y;
                    ai[j] += mr[j][cb] * c.y + mi[j][cb] * c.x;
                }
                // bounds how far the scheduler hoists loads (register pressure)
                if ((cb & 1) == 1) __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int j = 0; j < RBW; ++j)
                if (okj[j]) {
                    const int fl = flj[j], row = rowj[j];
                    T xr = sum_xor16(ar[j]), xi = sum_xor16(ai[j]);
                    xr = sum_xor32(xr);
                    xi = sum_xor32(xi);
                    if (kq == 0) s_z[fl * r16 + row] = mk2<C2>(xr, xi);
                    if (row == (fl ? cj : ci)) {
#pragma unroll
                        for (int cb = 0; cb < CB; ++cb) s_row[fl * NC + 4 * cb + kq] = mk2<C2>(mr[j][cb], mi[j][cb]);
                    }
                }
        }
        __syncthreads();
        // with the reduction's partials (its barrier also publishes alpha and the dets): one
        // thread per (state, flavour), the two flavours of state k in lanes 2k and 2k + 1;
        // the even lane takes its partner's determinant by DPP and forms the amplitude
        double ls = 0.0, bv = -1.0;
        int bk = nk;
        for (int e0 = 0; e0 < 2 * nk; e0 += NT) {
            const int e = e0 + tid, k = e >> 1, fl = e & 1;
            const bool on = e < 2 * nk;
            zc d = zmk(0.0, 0.0);
            if (on) {
                zc B[9];
                cap_move(k, fl, ci, sa, sb, cj, B, PF && e0 == 0, pf[0], pf[1]);
                d = det3(B);
                s_dets[e] = d;
            }
            const zc d1 = zmk(__longlong_as_double((long long)dpp_u64<0xB1>((unsigned long long)__double_as_longlong(d.x), 0ull)),
                              __longlong_as_double((long long)dpp_u64<0xB1>((unsigned long long)__double_as_longlong(d.y), 0ull)));
            if (on && fl == 0) {
                const zc tot = zmul(zmul(zmk((double)s_hsg[k], 0.0), d), d1);
                s_alpha[k] = tot;
                const double v = znorm(tot);
                ls += v;
                if (v > bv) {
                    bv = v;
                    bk = k;
                }
            }
        }
        reduce_partials(ls, bv, bk, S2, kn);
    };
    // move to state kn after eval0 (mv false) or eval_mv (mv true)
    auto apply = [&](bool mv, int ci, int sa, int sb, int cj, int kn) __attribute__((always_inline)) {
        const int m0 = mv ? 1 : 0;
        const int Rk0 = __builtin_amdgcn_readfirstlane(s_hn[kn]), Rk1 = __builtin_amdgcn_readfirstlane(s_hn[nk + kn]);
        auto RK = [&](int fl) { return fl ? Rk1 : Rk0; };
        // columns f_m of M and the Q rows of the filled orbitals, before any change
#pragma unroll
        for (int j = 0; j < RBW; ++j) {
            int fl, row;
            if (rinfo(j, fl, row)) {
#pragma unroll
                for (int m = 0; m < 2; ++m) {
                    if (m >= RK(fl)) break;
                    const int fm = s_hs[(fl * 2 + m) * nk + kn];
                    if (kq == (fm & 3)) {
                        T vr = 0, vi = 0;
#pragma unroll
                        for (int cb = 0; cb < CB; ++cb)
                            if (cb == (fm >> 2)) {
                                vr = mr[j][cb];
                                vi = mi[j][cb];
                            }
                        s_colb[(fl * 2 + m) * r16 + row] = mk2<C2>(vr, vi);
                    }
                    if (row == NF(fl) + s_hq[(fl * 2 + m) * nk + kn]) {
#pragma unroll
                        for (int cb = 0; cb < CB; ++cb)
                            s_vq[(fl * 2 + m) * NC + 4 * cb + kq] = mk2<C2>(mr[j][cb], mi[j][cb]);
                    }
                }
            }
        }
        __syncthreads();
        if (tid < 2) {
            const int fl = tid, R = RK(fl), D = R + m0;
            if (D > 0) {
                zc B[9];
                if (mv) {
                    cap_move(kn, fl, ci, sa, sb, cj, B);
                } else {
                    eye3(B);
#pragma unroll
                    for (int m = 0; m < 2; ++m)
#pragma unroll
                        for (int q = 0; q < 2; ++q)
                            if (m < R && q < R) B[m * 3 + q] = s_ra[((fl * 2 + m) * 2 + q) * nk + kn];
                }
                inv3(B, s_binv + fl * 9);
            }
        }
        // update coefficients of every row
        for (int e = tid; e < 2 * r16; e += NT) {
            const int fl = e >= r16 ? 1 : 0, i = e - fl * r16, R = RK(fl);
            const int r = fl ? cj : ci, snew = fl ? sa : sb, sold = fl ? sb : sa;
            zc kp[3] = {zmk(0, 0), zmk(0, 0), zmk(0, 0)}, dd = zmk(0, 0);
            int rp = -1;
            if (i < NF(fl)) {
                if (mv) {
                    kp[0] = zsub(toz(s_z[fl * r16 + i]), zmk(i == r ? 1.0 : 0.0, 0.0));
#pragma unroll
                    for (int m = 0; m < 2; ++m)
                        if (m < R) kp[1 + m] = toz(s_colb[(fl * 2 + m) * r16 + i]);
                } else {
#pragma unroll
                    for (int m = 0; m < 2; ++m)
                        if (m < R) kp[m] = toz(s_colb[(fl * 2 + m) * r16 + i]);
                }
            } else if (i < NF(fl) + NQ(fl)) {
                const int o = s_qorb[fl * nqm + (i - NF(fl))];
                if (o >= 0) {
                    int ms = -1;
#pragma unroll
                    for (int m = 0; m < 2; ++m)
                        if (m < R && s_hto[(fl * 2 + m) * nk + kn] == o) ms = m;
                    if (ms >= 0) {
                        const int o1 = s_hfr[(fl * 2 + ms) * nk + kn];
                        rp = s_slot[fl * Ns + o1];
                        if (mv) {
                            const zc eps = zsub(phi(fl, o1, snew), phi(fl, o1, sold));
                            dd = eps;
                            kp[0] = zmul(eps, toz(s_z[fl * r16 + r]));
#pragma unroll
                            for (int m = 0; m < 2; ++m)
                                if (m < R)
                                    kp[1 + m] = zadd(zmk(m == ms ? 1.0 : 0.0, 0.0),
                                                     zmul(eps, toz(s_row[fl * NC + s_hs[(fl * 2 + m) * nk + kn]])));
                        } else {
#pragma unroll
                            for (int m = 0; m < 2; ++m)
                                if (m < R) kp[m] = zmk(m == ms ? 1.0 : 0.0, 0.0);
                        }
                    } else if (mv) {
                        const zc pn = phi(fl, o, snew), dq = zsub(pn, phi(fl, o, sold));
                        dd = dq;
                        kp[0] = zadd(zsub(toz(s_z[fl * r16 + i]), pn), zmul(dq, toz(s_z[fl * r16 + r])));
#pragma unroll
                        for (int m = 0; m < 2; ++m)
                            if (m < R)
                                kp[1 + m] = zadd(toz(s_colb[(fl * 2 + m) * r16 + i]),
                                                 zmul(dq, toz(s_row[fl * NC + s_hs[(fl * 2 + m) * nk + kn]])));
                    } else {
#pragma unroll
                        for (int m = 0; m < 2; ++m)
                            if (m < R) kp[m] = toz(s_colb[(fl * 2 + m) * r16 + i]);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) s_kap[(fl * r16 + i) * 3 + q] = mk2<C2>(kp[q].x, kp[q].y);
            s_dl[fl * r16 + i] = mk2<C2>(dd.x, dd.y);
            s_rep[fl * r16 + i] = rp;
        }
        __syncthreads();
        // W = B^-1 V
        for (int e = tid; e < 2 * NC; e += NT) {
            const int fl = e >= NC ? 1 : 0, f = e - fl * NC, R = RK(fl), D = R + m0;
            zc V[3] = {zmk(0, 0), zmk(0, 0), zmk(0, 0)};
            if (mv) {
                V[0] = toz(s_row[fl * NC + f]);
#pragma unroll
                for (int m = 0; m < 2; ++m)
                    if (m < R) {
                        zc v = toz(s_vq[(fl * 2 + m) * NC + f]);
                        if (f == s_hs[(fl * 2 + m) * nk + kn]) v.x -= 1.0;
                        V[1 + m] = v;
                    }
            } else {
#pragma unroll
                for (int m = 0; m < 2; ++m)
                    if (m < R) {
                        zc v = toz(s_vq[(fl * 2 + m) * NC + f]);
                        if (f == s_hs[(fl * 2 + m) * nk + kn]) v.x -= 1.0;
                        V[m] = v;
                    }
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                zc acc = zmk(0.0, 0.0);
                if (q < D) {
#pragma unroll
                    for (int b = 0; b < 3; ++b) acc = zadd(acc, zmul(s_binv[fl * 9 + q * 3 + b], V[b]));
                }
                s_wr[(fl * 3 + q) * NC + f] = mk2<C2>(acc.x, acc.y);  // zero past D
            }
        }
        __syncthreads();
        // the update of every register row: column blocks outer, no branch inside (the
        // coefficients of rows that do not change, and the W rows past D, are zero)
        {
            C2 kj0[RBW], kj1[RBW], kj2[RBW], dj[RBW];
            int rpj[RBW], flj[RBW];
#pragma unroll
            for (int j = 0; j < RBW; ++j) {
                int fl = 0, row = 0;
                kj0[j] = kj1[j] = kj2[j] = dj[j] = mk2<C2>(0.0, 0.0);
                rpj[j] = -1;
                flj[j] = 0;
                if (rinfo(j, fl, row) && RK(fl) + m0 > 0) {
                    const int ix = fl * r16 + row;
                    flj[j] = fl;
                    kj0[j] = s_kap[ix * 3];
                    kj1[j] = s_kap[ix * 3 + 1];
                    kj2[j] = s_kap[ix * 3 + 2];
                    dj[j] = s_dl[ix];
                    rpj[j] = s_rep[ix];
                }
            }
            // four rank-1 passes (W_0, W_1, W_2, then + delta V_0), each loading one complex
            // value per element: the register pressure of one pass, not of four
            const int dmax = __builtin_amdgcn_readfirstlane(Rk0 > Rk1 ? Rk0 + m0 : Rk1 + m0);
#pragma unroll
            for (int pass = 0; pass < 4; ++pass) {
                // a scalar branch per pass: passes stay separate blocks, and unneeded ones
                // (W_a past every D, V_0 without a move) are skipped
                if (pass < 3 ? pass >= dmax : !mv) continue;
#pragma unroll
                for (int cb = 0; cb < CB; ++cb) {
                    const int f = 4 * cb + kq;
#pragma unroll
                    for (int j = 0; j < RBW; ++j) {
                        const C2 kc = pass == 0 ? kj0[j] : pass == 1 ? kj1[j] : pass == 2 ? kj2[j] : dj[j];
                        const C2 wv2 = pass < 3 ? s_wr[(flj[j] * 3 + pass) * NC + f] : s_row[flj[j] * NC + f];
                        T br = mr[j][cb], bi = mi[j][cb];
                        if (pass == 0 && rpj[j] >= 0) {
                            br = f == rpj[j] ? (T)1 : (T)0;
                            bi = 0;
                        }
                        const T sg = pass < 3 ? (T)-1 : (T)1;
                        br += sg * (kc.x * wv2.x - kc.y * wv2.y);
                        bi += sg * (kc.x * wv2.y + kc.y * wv2.x);
                        mr[j][cb] = br;
                        mi[j][cb] = bi;
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
                asm volatile("" ::: "memory");
            }
        }
        // amplitude and slot / Q-row bookkeeping, the independent parts side by side: one lane
        // per hop (fl, m) in wavefront 0 (a state's hops empty and fill distinct orbitals, so
        // they touch distinct slots and Q rows), one per flavour's amplitude in wavefront 1,
        // the site map in wavefront 2
        if (wv == 0 && lane < 4) {
            const int fl = lane >> 1, m = lane & 1;
            if (m < RK(fl)) {
                const int o1 = s_hfr[(fl * 2 + m) * nk + kn], o2 = s_hto[(fl * 2 + m) * nk + kn];
                const int fm = s_slot[fl * Ns + o1];
                s_slot[fl * Ns + o1] = -1;
                s_slot[fl * Ns + o2] = fm;
                s_cur[fl * nmax + fm] = o2;
                const int q = s_qrow[fl * Ns + o2];
                s_qrow[fl * Ns + o2] = -1;
                s_qrow[fl * Ns + o1] = q;
                s_qorb[fl * nqm + q] = o1;
            }
        }
        if (wv == (NW > 1 ? 1 : 0) && lane >= 4 && lane < 6) {
            const int fl = lane - 4;
            const zc d = s_dets[kn * 2 + fl];
            zamp_mul(s_amp[2 * fl], s_amp[2 * fl + 1], s_amp[4 + fl], d.x, d.y);
        }
        if (mv && wv == (NW > 2 ? 2 : 0) && lane == 6) {
            s_site[sb] = ci;
            s_site[sa] = (1 << 16) | cj;
            s_part[ci] = sb;
            s_part[nmax + cj] = sa;
        }
        // the hop graph's row of the current state changes only with the state (the slot
        // and Q-row bookkeeping above moves nothing when kn is the current state)
        const bool moved_state = kn != cs;
        cs = kn;
        __syncthreads();
        if (moved_state) cache_hops();
        build_ra();
    };

    if (tid == 0) {
        s_amp[0] = 1.0; s_amp[1] = 0.0; s_amp[2] = 1.0; s_amp[3] = 0.0;
        s_amp[4] = 0.0; s_amp[5] = 0.0;
    }
    cache_hops();
    build_ra();
    double S;
    int kmax;
    eval0();
    reduce_alpha(S, kmax);
    // mode 2: the hop to the dominant state runs as step -1 of the walk loop (no lattice
    // move), so that the kernel holds a single copy of apply
    const bool hop0 = MK == 0 && a.mode == 2 && kmax != cs && kmax < nk;
    int64_t n_noop = 0, n_acc = 0, n_rej = 0;
    double sacc[2] = {0, 0};  // n_meas, sum ln w
    const uint64_t ctr0 = a.counter[w];
    int64_t mslot = 0;
    C2 *Zg = a.zscr ? reinterpret_cast<C2 *>(a.zscr) + (size_t)w * z.zt * RS : nullptr;
    // ProjHeis::measure on the channel (ProjHeis.cpp:40-154)
    auto measure = [&](bool acc_on) __attribute__((always_inline)) {
        if constexpr (MK == 0) return;
        eval0();
        double S0;
        int kk;
        reduce_alpha(S0, kk);
        for (int k = tid; k < nk; k += NT) s_a0[k] = s_alpha[k];
        // (the bond sums below read [Ai; Q] from M in memory: a measurement launch loaded
        // its registers from there and changes nothing)
        // Z_fl = M_fl x Phi_fl(current orbitals, sites of the other flavour's particles),
        // 16 targets per tile, on MFMA: Zg[target][row] (MK = 1), or in LDS (MK = 2) only the
        // entries the bond sums read: the Q rows, Zq_fl[target][q], and of the inverse rows
        // the entry y0[fl][t][d] of target t and the flavour-fl particle on its bond d
        C2 *Ct = reinterpret_cast<C2 *>(s_un);
        C2 *const s_zq = reinterpret_cast<C2 *>(smem + lo.mz);
        C2 *const s_y0 = s_zq + (size_t)z.zt * (z.nq0 + z.nq1);
        int *const s_nbt = reinterpret_cast<int *>(s_y0 + 8 * nmax);
        auto zq_at = [&](int fl, int tg, int q) __attribute__((always_inline)) {
            return s_zq + (fl ? (size_t)z.zt * z.nq0 : 0) + (size_t)tg * NQ(fl) + q;
        };
        if constexpr (LZ) {
            // nbt[fl][t]: byte d holds the flavour-fl particle on bond d of target t (a
            // particle of the other flavour), 0xFF where that site holds flavour 1 - fl
            // (published by the barrier after the first tile is staged)
            for (int i = tid; i < 2 * nmax; i += NT) {
                const int fl = i >= nmax ? 1 : 0, t = i - fl * nmax;
                unsigned pk = 0xFFFFFFFFu;
                if (t < NF(1 - fl)) {
                    const int st = s_part[(1 - fl) * nmax + t];
#pragma unroll
                    for (int d = 0; d < 4; ++d) {
                        const int v = s_site[s_nbr[st * 4 + d]];
                        if ((v >> 16) == fl) pk = (pk & ~(0xFFu << (8 * d))) | ((unsigned)(v & 0xFF) << (8 * d));
                    }
                }
                s_nbt[i] = (int)pk;
            }
        }
        // the tiles (flavour, first target) in sequence; each tile's Phi is loaded into
        // registers while the previous tile's MFMAs run, then stored to LDS
        constexpr int SPT = (NC * 16 + NT - 1) / NT;
        C2 sv[SPT];
        auto stage_ld = [&](int fl, int t0) __attribute__((always_inline)) {
            const int no = NF(1 - fl);
#pragma unroll
            for (int i = 0; i < SPT; ++i) {
                const int e = tid + i * NT, f = e >> 4, tg = t0 + (e & 15);
                C2 v = mk2<C2>(0.0, 0.0);
                if (e < NC * 16 && f < NF(fl) && tg < no)
                    v = tabT[((size_t)fl * Ns + s_part[(1 - fl) * nmax + tg]) * Ns + s_cur[fl * nmax + f]];
                sv[i] = v;
            }
        };
        stage_ld(0, 0);
        for (int fl = 0; fl < 2; ++fl) {
            const int no = NF(1 - fl);
            for (int t0 = 0; t0 < no; t0 += 16) {
#pragma unroll
                for (int i = 0; i < SPT; ++i)
                    if (tid + i * NT < NC * 16) Ct[tid + i * NT] = sv[i];
                __syncthreads();
                {
                    const bool last = t0 + 16 >= no;
                    if (!last) stage_ld(fl, t0 + 16);
                    else if (fl == 0) stage_ld(1, 0);
                }
#pragma unroll
                for (int j = 0; j < RBW; ++j) {
                    int f_, row;
                    if (rinfo(j, f_, row) && f_ == fl) {
                        typename MF::v4 zr = {0, 0, 0, 0}, zi = {0, 0, 0, 0};
#pragma unroll
                        for (int cb = 0; cb < CB; ++cb) {
                            const C2 b = Ct[(4 * cb + kq) * 16 + li];
                            zr = MF::mma(mr[j][cb], b.x, zr);
                            zr = MF::mma(mi[j][cb], -b.y, zr);
                            zi = MF::mma(mr[j][cb], b.y, zi);
                            zi = MF::mma(mi[j][cb], b.x, zi);
                        }
                        const int rbg = wv + j * NW;
                        if constexpr (LZ) {
                            const int tg = t0 + li, rb_f = rbg - (fl ? RB0 : 0);
                            const unsigned nb = tg < no ? (unsigned)s_nbt[fl * nmax + tg] : 0xFFFFFFFFu;
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const int rw = rb_f * 16 + MF::orow(lane, v);
                                const C2 val = mk2<C2>(zr[v], zi[v]);
                                if (tg < no) {
                                    if (rw < NF(fl)) {
#pragma unroll
                                        for (int d = 0; d < 4; ++d)
                                            if (((nb >> (8 * d)) & 0xFFu) == (unsigned)rw)
                                                s_y0[(fl * nmax + tg) * 4 + d] = val;
                                    } else if (rw < NF(fl) + NQ(fl)) {
                                        *zq_at(fl, tg, rw - NF(fl)) = val;
                                    }
                                }
                            }
                        } else {
#pragma unroll
                            for (int v = 0; v < 4; ++v)
                                Zg[(size_t)(t0 + li) * RS + rbg * 16 + MF::orow(lane, v)] = mk2<C2>(zr[v], zi[v]);
                        }
                    }
                }
                __syncthreads();
            }
        }
        // antiparallel bonds in bond order (a ballot prefix sum per round of NT bonds) in the
        // area the contraction tiles used: bl[3 b] = ci | cj << 16, bl[3 b + 1] = sa | sb << 16,
        // bl[3 b + 2] = d0 | d1 << 8 (up particle ci at site sa, down particle cj at site sb,
        // sb = nbr[sa][d0], sa = nbr[sb][d1])
        int *const bl = reinterpret_cast<int *>(s_un);
        int nb = 0;
        for (int e0 = 0; e0 < 2 * Ns; e0 += NT) {
            const int e = e0 + tid;
            bool ap = false;
            int v0 = 0, v1 = 0, v2 = 0;
            if (e < 2 * Ns) {
                const int4 eg = a.edge[e];
                const int ii = s_site[eg.x], ij = s_site[eg.z];
                ap = (ii >> 16) != (ij >> 16);
                if ((ii >> 16) == 0) {
                    v0 = (ii & 0xFFFF) | ((ij & 0xFFFF) << 16);
                    v1 = eg.x | (eg.z << 16);
                    v2 = eg.y | (eg.w << 8);
                } else {
                    v0 = (ij & 0xFFFF) | ((ii & 0xFFFF) << 16);
                    v1 = eg.z | (eg.x << 16);
                    v2 = eg.w | (eg.y << 8);
                }
            }
            const unsigned long long bm = __ballot(ap);
            const int pre = __builtin_amdgcn_mbcnt_hi((unsigned)(bm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bm, 0u));
            if (lane == 0) s_ired[wv] = __popcll(bm);
            __syncthreads();
            int off = nb, tot = 0;
            for (int q = 0; q < NW; ++q) {
                if (q < wv) off += s_ired[q];
                tot += s_ired[q];
            }
            if (ap) {
                bl[3 * (off + pre)] = v0;
                bl[3 * (off + pre) + 1] = v1;
                bl[3 * (off + pre) + 2] = v2;
            }
            nb += tot;
            __syncthreads();
        }
        // every antiparallel bond's ratio of every state: a thread per (state, group of
        // bonds), every load issued before the hop counts select (the cached hop data holds
        // valid indices past the hop count), the next bond's loads in flight
        const int ngrp = z.ngrp;
        for (int e2 = tid; e2 < nk * ngrp; e2 += NT) {
            const int k = e2 % nk, gq = e2 / nk;
            struct BondIn {
                C2 y0[2], aq[2][2], kq[2][2];
                zc p2s[2][2];
            };
            auto bld = [&](int b, BondIn &in) __attribute__((always_inline)) {
                const int bc = b < nb ? b : nb - 1;
                const int v0 = bl[3 * bc], v1 = bl[3 * bc + 1], v2 = bl[3 * bc + 2];
                const int ci = v0 & 0xFFFF, cj = v0 >> 16, sa = v1 & 0xFFFF, sb = v1 >> 16;
#pragma unroll
                for (int fl = 0; fl < 2; ++fl) {
                    const int r = fl ? cj : ci, tg = fl ? ci : cj;
                    const int snew = fl ? sa : sb, base = fl ? RB0 * 16 : 0;
                    const C2 *zcol = Zg + (size_t)tg * RS + base;
                    if constexpr (LZ)
                        in.y0[fl] = s_y0[(fl * nmax + tg) * 4 + ((v2 >> (fl ? 0 : 8)) & 3)];
                    else
                        in.y0[fl] = zcol[r];
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        in.aq[fl][m] = Mg[(size_t)(base + r) * NC + s_hs[(fl * 2 + m) * nk + k]];
                        const int o2 = s_hto[(fl * 2 + m) * nk + k];
                        in.p2s[fl][m] = phi(fl, o2 >= 0 ? o2 : 0, snew);
                        if constexpr (LZ)
                            in.kq[fl][m] = *zq_at(fl, tg, s_hq[(fl * 2 + m) * nk + k]);
                        else
                            in.kq[fl][m] = zcol[NF(fl) + s_hq[(fl * 2 + m) * nk + k]];
                    }
                }
            };
            zc hs = zmk(0.0, 0.0);
            auto bfin = [&](const BondIn &in) __attribute__((always_inline)) {
                zc tot = zmk((double)s_hsg[k], 0.0);
#pragma unroll
                for (int fl = 0; fl < 2; ++fl) {
                    const int R = s_hn[fl * nk + k];
                    zc B[9];
                    eye3(B);
                    B[0] = toz(in.y0[fl]);
#pragma unroll
                    for (int q = 0; q < 2; ++q)
                        if (q < R) B[1 + q] = toz(in.aq[fl][q]);
#pragma unroll
                    for (int m = 0; m < 2; ++m)
                        if (m < R) {
                            B[(1 + m) * 3] = zsub(toz(in.kq[fl][m]), in.p2s[fl][m]);
#pragma unroll
                            for (int q = 0; q < 2; ++q)
                                if (q < R) B[(1 + m) * 3 + 1 + q] = s_ra[((fl * 2 + m) * 2 + q) * nk + k];
                        }
                    tot = zmul(tot, det3(B));
                }
                hs = zadd(hs, tot);
            };
            if (gq < nb) {
                BondIn ia, ib;
                bld(gq, ia);
                for (int b = gq; b < nb; b += 2 * ngrp) {
                    bld(b + ngrp, ib);
                    bfin(ia);
                    bld(b + 2 * ngrp, ia);
                    if (b + ngrp < nb) bfin(ib);
                }
            }
            s_hpart[gq * nk + k] = hs;
        }
        if (tid == 0) s_ired[15] = 0;
        __syncthreads();
        {
            int ap = 0;
            for (int e = tid; e < 2 * Ns; e += NT) {
                const int4 eg = a.edge[e];
                ap += (s_site[eg.x] >> 16) != (s_site[eg.z] >> 16);
            }
            atomicAdd(&s_ired[15], ap);
        }
        __syncthreads();
        const int nap = s_ired[15];
        const double cpar = 0.25 * (double)(2 * Ns - nap) - 0.25 * (double)nap;
        for (int k = tid; k < nk; k += NT) {
            zc h = zmk(0.0, 0.0);
            for (int q = 0; q < ngrp; ++q) h = zadd(h, s_hpart[q * nk + k]);
            const zc a0 = s_a0[k];
            s_alpha[k] = zmk(-0.5 * h.x + cpar * a0.x, -0.5 * h.y + cpar * a0.y);  // h_k
        }
        __syncthreads();
        const double lw = 2.0 * (a.logabs[2 * w] + a.logabs[2 * w + 1]) + log(S0);
        if (acc_on) {
            // the record of this measurement; excr_fold_kernel adds it to the H and O sums
            double2 *hb = a.hbuf + ((size_t)w * a.hb_cap + a.hb_slot) * (2 * nk + 1);
            for (int k = tid; k < nk; k += NT) {
                const zc a0 = s_a0[k], hk = s_alpha[k];
                hb[k] = make_double2(a0.x, a0.y);
                hb[nk + k] = make_double2(hk.x, hk.y);
            }
            if (tid == 0) {
                hb[2 * nk] = make_double2(S0, 0.0);
                sacc[0] += 1;
                sacc[1] += lw;
            }
        }
        if (tid == 0 && a.tr_logw) a.tr_logw[(size_t)w * a.n_meas_slots + a.meas_slot + mslot] = lw;
        ++mslot;
        __syncthreads();
    };

    if constexpr (MK != 0) {
        measure(a.mode == 5);
    } else if (a.mode == 4) {
        const double la = a.logabs[2 * w] + a.logabs[2 * w + 1], ph = a.phase[2 * w] + a.phase[2 * w + 1];
        for (int k = tid; k < nk; k += NT) {
            const zc al = s_alpha[k];
            const double nrm = znorm(al);
            a.amp_out[((size_t)w * nk + k) * 2] = nrm > 0.0 ? la + 0.5 * log(nrm) : -INFINITY;
            a.amp_out[((size_t)w * nk + k) * 2 + 1] = remainder(ph + atan2(al.y, al.x), 2 * M_PI);
        }
    } else if (MK == 0 && (a.mode == 0 || hop0)) {
        const int64_t n_walk = a.mode == 0 ? a.n_steps : 0;
        // Philox draws of 64 steps at a time, lane l holding step (st & ~63) + l: the site,
        // the bond and the acceptance word
        int bv1 = 0, bed = 0;
        uint32_t bc3 = 0;
        for (int64_t st = hop0 ? -1 : 0; st < n_walk; ++st) {
            asm volatile("" : "+v"(tid), "+v"(lane), "+v"(li), "+v"(kq));
            bool mv = st >= 0, acc = true, noop = false;
            int ci = 0, cj = 0, sa = 0, sb = 0, kn = kmax;
            double ratio = 1.0, S2 = 0.0;
            if (mv) {
                int v1, ed;
                double u3;
                if (a.uniforms) {
                    const double *up = a.uniforms + ((size_t)w * a.n_total + a.st0 + st) * 4;
                    v1 = (int)(Ns * up[0]);
                    v1 = v1 < Ns ? v1 : Ns - 1;
                    ed = (int)(4 * up[1]);
                    ed = ed < 4 ? ed : 3;
                    u3 = up[3];
                } else {
                    if ((st & 63) == 0) {
                        const uint64_t sc = ctr0 + (uint64_t)(st + lane);
                        uint32_t cc[4] = {(uint32_t)sc, (uint32_t)(sc >> 32), (uint32_t)w, 0u};
                        gv::philox4x32_10_dev(cc, a.key0, a.key1);
                        bv1 = (int)(((uint64_t)Ns * (2ull * cc[0] + 1ull)) >> 33);
                        bed = (int)(cc[1] >> 30);
                        bc3 = cc[3];
                    }
                    const int l = (int)(st & 63);
                    v1 = __builtin_amdgcn_readlane(bv1, l);
                    ed = __builtin_amdgcn_readlane(bed, l);
                    u3 = ((double)(uint32_t)__builtin_amdgcn_readlane((int)bc3, l) + 0.5) * 2.3283064365386963e-10;
                }
                v1 = __builtin_amdgcn_readfirstlane(v1);
                ed = __builtin_amdgcn_readfirstlane(ed);
                const int v2 = __builtin_amdgcn_readfirstlane(s_nbr[v1 * 4 + ed]);
                const int i1 = __builtin_amdgcn_readfirstlane(s_site[v1]), i2 = __builtin_amdgcn_readfirstlane(s_site[v2]);
                noop = (i1 >> 16) == (i2 >> 16);
                if (noop) {
                    acc = u3 < 1.0;
                    ++n_noop;
                } else {
                    if ((i1 >> 16) == 0) {
                        ci = i1 & 0xFFFF;
                        sa = v1;
                        cj = i2 & 0xFFFF;
                        sb = v2;
                    } else {
                        ci = i2 & 0xFFFF;
                        sa = v2;
                        cj = i1 & 0xFFFF;
                        sb = v1;
                    }
                    eval_mv(ci, sa, sb, cj, S2, kn);
                    ratio = S2 / S;
                    acc = ratio > 1.0 ? (u3 < 1.0) : (u3 < ratio);  // MetroMC::YesNo
                    acc = __builtin_amdgcn_readfirstlane(acc ? 1 : 0) != 0;
                }
            }
            if (!noop && acc) {
                const zc ak = s_alpha[kn];
                apply(mv, ci, sa, sb, cj, kn);
                if (mv) {
                    S = S2 / znorm(ak);
                    ++n_acc;
                }
            } else if (!noop) {
                ++n_rej;
            }
            if (mv && tid == 0) {
                const size_t ti = (size_t)w * a.n_total + a.st0 + st;
                if (a.tr_noop) a.tr_noop[ti] = noop ? 1 : 0;
                if (a.tr_accept) a.tr_accept[ti] = acc ? 1 : 0;
                if (a.tr_ratio) a.tr_ratio[ti] = ratio;
            }
        }
        if (hop0) {
            eval0();
            reduce_alpha(S, kmax);
        }
        // only a walk changes [Ai; Q]: storing it on this branch alone ends the registers'
        // live range here, so the other branches' work does not add to the walk's peak
        store_regs();
    }
    // write back the configuration (a measurement launch changes neither it nor [Ai; Q])
    __syncthreads();
    if constexpr (MK == 0) {
        for (int i = tid; i < Ns; i += NT) a.site[(size_t)w * Ns + i] = s_site[i];
        for (int i = tid; i < 2 * nmax; i += NT) {
            a.part[(size_t)w * 2 * nmax + i] = s_part[i];
            a.cur[(size_t)w * 2 * nmax + i] = s_cur[i];
        }
        for (int i = tid; i < 2 * nqm; i += NT) a.qorb[(size_t)w * 2 * nqm + i] = s_qorb[i];
    }
    if (tid == 0) {
        if (MK == 0) {
            // ln|det| and arg det per flavour += the launch's product of ratios
#pragma unroll
            for (int fl = 0; fl < 2; ++fl) {
                const double mr = s_amp[2 * fl], mi = s_amp[2 * fl + 1];
                a.logabs[2 * w + fl] += log(hypot(mr, mi)) + s_amp[4 + fl] * 0.69314718055994531;
                a.phase[2 * w + fl] = remainder(a.phase[2 * w + fl] + atan2(mi, mr), 2 * M_PI);
            }
        }
        a.cstate[w] = cs;
        a.weight[w] = S;
        if (MK != 0 && a.mode == 5) {
            double *ac = a.acc + (size_t)w * gv::ACC_STRIDE;
            ac[3] += sacc[0];
            ac[11] += sacc[1];
        }
        if (MK == 0 && a.mode == 0) {
            a.counter[w] = ctr0 + (uint64_t)a.n_steps;
            double *ac = a.acc + (size_t)w * gv::ACC_STRIDE;
            ac[7] += (double)a.n_steps;
            ac[8] += (double)n_acc;
            ac[9] += (double)n_rej;
            ac[10] += (double)n_noop;
        }
    }
}

// M of every walker from the FP64 inverse of its current state (after the Gauss-Jordan
// of a refresh): the Ai rows, then Q[q][f] = sum_r Phi(o_q, site r) Ai[r][f] in FP64 for
// the new orbitals of the current state in increasing order.  One block per
// (walker, flavour).
template <typename T> __global__ void __launch_bounds__(256) excr_mbuild_kernel(ExcArgs a, const double2 *Ai64)
{
    using C2 = typename cpx<T>::t;
    const int w = blockIdx.x >> 1, fl = blockIdx.x & 1;
    const int Ns = a.L * a.L, nk = a.nk, NC = 4 * a.cb;
    const int n = fl ? a.n_dn : a.n_up, nq = fl ? a.nq1 : a.nq0;
    const int nmax = a.n_up > a.n_dn ? a.n_up : a.n_dn, nqm = a.nq0 > a.nq1 ? a.nq0 : a.nq1;
    const int RBT = a.rb0 + a.rb1, rows = 16 * (fl ? a.rb1 : a.rb0);
    const size_t chb = (size_t)(a.chan ? a.chan[w] : 0);
    const int c = a.cstate[w];
    const int32_t *off = a.new_off + chb * (2 * nk + 1) + c * 2 + fl;
    const int32_t *no = a.new_orb + off[0];
    const int nnew = off[1] - off[0];
    const double2 *Ai = Ai64 + ((size_t)w * 2 + fl) * nmax * nmax;
    const double2 *tab = reinterpret_cast<const double2 *>(a.tab64) + ((size_t)a.group[w] * 2 + fl) * Ns * Ns;
    const int32_t *part = a.part + (size_t)w * 2 * nmax + (size_t)fl * nmax;
    C2 *M = reinterpret_cast<C2 *>(a.M) + ((size_t)w * RBT * 16 + (fl ? a.rb0 * 16 : 0)) * NC;
    for (int e = threadIdx.x; e < rows * NC; e += blockDim.x) {
        const int row = e / NC, f = e - row * NC;
        double x = 0.0, y = 0.0;
        if (f < n) {
            if (row < n) {
                const double2 v = Ai[(size_t)row * nmax + f];
                x = v.x;
                y = v.y;
            } else if (row < n + nq && row - n < nnew) {
                const int o = no[row - n];
                for (int r = 0; r < n; ++r) {
                    const double2 p = tab[(size_t)o * Ns + part[r]], q = Ai[(size_t)r * nmax + f];
                    x += p.x * q.x - p.y * q.y;
                    y += p.x * q.y + p.y * q.x;
                }
            }
        }
        M[(size_t)row * NC + f] = mk2<C2>(x, y);
    }
    for (int q = threadIdx.x; q < nqm; q += blockDim.x)
        a.qorb[((size_t)w * 2 + fl) * nqm + q] = (q < nq && q < nnew) ? no[q] : -1;
}

// ProjHeis H and O sums (ProjHeis.cpp:140-153) of the buffered measurements, in measurement
// order: H[k1][k2] += conj(a_k1) h_k2 / S, O[k1][k2] += conj(a_k1) a_k2 / S.  One block per
// walker; the records are staged in LDS and each sum is loaded and stored once per fold
// instead of once per measurement.
__global__ void __launch_bounds__(256) excr_fold_kernel(ExcArgs a, int cnt)
{
    extern __shared__ double2 s_rec[];
    const int w = blockIdx.x, nk = a.nk, per = 2 * nk + 1;
    const double2 *hb = a.hbuf + (size_t)w * a.hb_cap * per;
    for (int i = threadIdx.x; i < cnt * per; i += blockDim.x) s_rec[i] = hb[i];
    __syncthreads();
    double2 *HO = reinterpret_cast<double2 *>(a.ho) + (size_t)w * 2 * nk * nk;
    const int ne = nk * nk;
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        const int k1 = e / nk, k2 = e - k1 * nk;
        double2 hs = HO[e], os = HO[ne + e];
        for (int m = 0; m < cnt; ++m) {
            const double2 *r = s_rec + m * per;
            const double S0 = r[2 * nk].x;
            const zc a1 = zmk(r[k1].x, r[k1].y), h2 = zmk(r[nk + k2].x, r[nk + k2].y), a2 = zmk(r[k2].x, r[k2].y);
            const zc hv = zconjmul(a1, h2), ov = zconjmul(a1, a2);  // conj(a_k1) h_k2, conj(a_k1) a_k2
            hs.x += hv.x / S0;
            hs.y += hv.y / S0;
            os.x += ov.x / S0;
            os.y += ov.y / S0;
        }
        HO[e] = hs;
        HO[ne + e] = os;
    }
}

struct Shape {
    int prec, cb, rbw;
    void (*kern)(ExcArgs);
    void (*meas)(ExcArgs);
    void (*meas_lds)(ExcArgs);
};

#define GR_SHAPE(P, TT, C, R) \
    {P, C, R, excr_kernel<TT, C, R, 0>, excr_kernel<TT, C, R, 1>, excr_kernel<TT, C, R, 2>}
const Shape kShapes[] = {
    GR_SHAPE(1, float, 3, 1),  GR_SHAPE(1, float, 5, 1),   GR_SHAPE(1, float, 9, 2),
    GR_SHAPE(1, float, 13, 2), GR_SHAPE(1, float, 19, 2),  GR_SHAPE(1, float, 19, 3),
    GR_SHAPE(0, double, 3, 1), GR_SHAPE(0, double, 5, 1),  GR_SHAPE(0, double, 9, 2),
    GR_SHAPE(0, double, 13, 2),
};
#undef GR_SHAPE

const Shape *find_shape(int prec, int cb, int rbw)
{
    for (const Shape &sh : kShapes)
        if (sh.prec == prec && sh.cb == cb && sh.rbw == rbw) return &sh;
    return nullptr;
}

}  // namespace gr

extern "C" int gv_excr_shape(int prec, int cb_min, int row_blocks, int *cb, int *rbw)
{
    for (const gr::Shape &sh : gr::kShapes)
        if (sh.prec == prec && sh.cb >= cb_min && sh.rbw * gr::NW >= row_blocks) {
            *cb = sh.cb;
            *rbw = sh.rbw;
            return 1;
        }
    return 0;
}

extern "C" size_t gv_excr_lds_bytes(int prec, int L, int n_up, int n_dn, int nq0, int nq1, int rb0, int rb1, int cb,
                                    int nk)
{
    const gr::Sizes z = gr::make_sizes(L, n_up, n_dn, nq0, nq1, rb0, rb1, cb, nk);
    return gr::carve(z, prec == 0 ? 16 : 8).total + 64;
}

extern "C" size_t gv_excr_meas_lds_bytes(int prec, int L, int n_up, int n_dn, int nq0, int nq1, int rb0, int rb1,
                                         int cb, int nk)
{
    const gr::Sizes z = gr::make_sizes(L, n_up, n_dn, nq0, nq1, rb0, rb1, cb, nk);
    return gr::carve(z, prec == 0 ? 16 : 8, true).total + 64;
}

extern "C" hipError_t gv_launch_excr(int prec, int rbw, const gv::ExcArgs *a, int nblocks, hipStream_t st)
{
    const gr::Shape *sh = gr::find_shape(prec, a->cb, rbw);
    if (!sh || (a->rb0 + a->rb1) > rbw * gr::NW) return hipErrorInvalidValue;
    const size_t lds = gv_excr_lds_bytes(prec, a->L, a->n_up, a->n_dn, a->nq0, a->nq1, a->rb0, a->rb1, a->cb, a->nk);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (a->mode == 1 || a->mode == 5) {
        // measurement: tiles in LDS where they fit, else in the HBM scratch
        const size_t lz =
            gv_excr_meas_lds_bytes(prec, a->L, a->n_up, a->n_dn, a->nq0, a->nq1, a->rb0, a->rb1, a->cb, a->nk);
        if (lz <= 160 * 1024) {
            hipLaunchKernelGGL(sh->meas_lds, dim3(nblocks), dim3(gr::NT), lz, st, *a);
        } else {
            if (!a->zscr) return hipErrorInvalidValue;
            hipLaunchKernelGGL(sh->meas, dim3(nblocks), dim3(gr::NT), lds, st, *a);
        }
        return hipGetLastError();
    }
    hipLaunchKernelGGL(sh->kern, dim3(nblocks), dim3(gr::NT), lds, st, *a);
    return hipGetLastError();
}

extern "C" hipError_t gv_launch_excr_fold(const gv::ExcArgs *a, int cnt, int nwalkers, hipStream_t st)
{
    if (cnt < 1 || cnt > a->hb_cap || !a->hbuf) return hipErrorInvalidValue;
    const size_t lds = (size_t)cnt * (2 * a->nk + 1) * sizeof(double2);
    if (lds > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gr::excr_fold_kernel, dim3(nwalkers), dim3(256), lds, st, *a, cnt);
    return hipGetLastError();
}

extern "C" hipError_t gv_launch_excr_mbuild(int prec, const gv::ExcArgs *a, const double2 *Ai64, int nwalkers,
                                            hipStream_t st)
{
    if (prec == 0)
        hipLaunchKernelGGL(gr::excr_mbuild_kernel<double>, dim3(2 * nwalkers), dim3(256), 0, st, *a, Ai64);
    else
        hipLaunchKernelGGL(gr::excr_mbuild_kernel<float>, dim3(2 * nwalkers), dim3(256), 0, st, *a, Ai64);
    return hipGetLastError();
}
