// FP64 inversion with log-determinant of batches of complex matrices: the role of
// linalg::DetInv (linalg.cpp:9-44) at init and at every drift refresh.
//
// Algorithm: Gauss-Jordan elimination with implicit partial pivoting.  Column k is
// eliminated with the pivot row p = the unpivoted row of largest |re| + |im| in column k
// (lowest row on ties; the zgetrf choice of linalg.cpp:16 applied to the unpivoted rows).
// The pivot row's own scaling is deferred: at its step the row keeps its values except a
// -1 in column k, and records the scale s_p = -1/pivot.  Every later step then treats it
// like any other row (U <- U - f r with f = U[k]/pivot), and the true row is s_p U at the
// end.  The update of a step is therefore one multiply-add per element with no branch:
//     U[i][j] -= f_i r_j   (f_i = U[i][k]/pivot, 0 for the pivot row),   U[i][k] = f_i
// det = sign(permutation) * prod(pivots); ln|det| and arg det are sums over the pivots.
// The matrix counts as singular when |det|^2 < 1e-9 (linalg.cpp:23) or a pivot is zero.
//
// Two kernels:
//  * di_reg_kernel<NP, LPR>: the working matrix in VGPRs, padded to NP x NP with the
//    identity (padding rows and columns never mix with the real block: their entries in
//    the real columns are exact zeros), thread (row, c) holding the LPR-strided columns.
//    Two barriers per column: the waves' pivot candidates (key and value) go through LDS,
//    then the winning row alone writes itself for the update.
//  * di_stream_kernel<NB>: matrices too large for a CU's registers (n > 128).  One row
//    per thread; the working matrix lives column-major in a work buffer, and columns are
//    eliminated in panels of NB: the panel's columns stay in registers during its NB
//    steps, each pivot row's full row is rebuilt from the panel-start values and the
//    earlier pivot rows of the panel, and the rest of every row is updated once per
//    panel (rank-NB), so the work buffer is read and written n/NB times instead of n.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "../../include/gpvmc.h"
#include "vmc_kernels.h"

namespace gd {

using C = double2;

__device__ __forceinline__ C cm(double x, double y)
{
    C r;
    r.x = x;
    r.y = y;
    return r;
}

// explicit multiply-adds, so that every kernel (and every inlining of it) rounds alike
__device__ __forceinline__ C cmul(C a, C b) { return cm(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x)); }

// acc - f r
__device__ __forceinline__ C cfms(C f, C r, C acc)
{
    acc.x = fma(-f.x, r.x, acc.x);
    acc.x = fma(f.y, r.y, acc.x);
    acc.y = fma(-f.x, r.y, acc.y);
    acc.y = fma(-f.y, r.x, acc.y);
    return acc;
}

// Block -> (matrix size, output slot) and the element (row, col) of the matrix before
// elimination.  Modes 0/1: block 2w + fl is flavour fl of walker w, row = slot f,
// col = label r, A(f, r) = Phi(orbital of slot f, site of label r).  Mode 2: block b is
// caller matrix b, column-major, A(row, col) at A[b][row + col n].
struct Blk {
    int n, w, fl, o;
    const double2 *base;
};

__device__ __forceinline__ Blk blk_of(const gv::DIArgs &a)
{
    Blk b;
    if (a.mode == 2) {
        b.n = a.n_up;
        b.w = blockIdx.x;
        b.fl = 0;
        b.o = blockIdx.x;
        b.base = a.A + (size_t)blockIdx.x * a.n_up * a.n_up;
    } else {
        b.w = blockIdx.x >> 1;
        b.fl = blockIdx.x & 1;
        b.n = b.fl ? a.n_dn : a.n_up;
        b.o = blockIdx.x;
        const size_t g2 = (size_t)a.group[b.w] * 2 + b.fl;
        b.base = a.mode == 0 ? a.occ + g2 * a.Ns * b.n : a.tab + g2 * a.Ns * a.Ns;
    }
    return b;
}

// s_col[r] = site of label r (modes 0/1), s_row[f] = orbital of slot f (mode 1)
__device__ __forceinline__ C elem(const gv::DIArgs &a, const Blk &b, const int *s_row, const int *s_col, int row,
                                  int col)
{
    if (a.mode == 0) return b.base[(size_t)s_col[col] * b.n + row];
    if (a.mode == 1) return b.base[(size_t)s_row[row] * a.Ns + s_col[col]];
    return b.base[(size_t)row + (size_t)col * b.n];
}

__device__ __forceinline__ void put(const gv::DIArgs &a, const Blk &b, int x, int y, C v)
{
    if (a.mode == 2) {
        a.Ai64[(size_t)b.o * b.n * b.n + (size_t)x + (size_t)y * b.n] = v;
    } else {
        const size_t i = ((size_t)b.o * a.nstride + x) * a.nstride + y;
        a.Ai64[i] = v;
        if (a.Ai32) a.Ai32[i] = make_float2((float)v.x, (float)v.y);
    }
}

__device__ __forceinline__ void stage_indices(const gv::DIArgs &a, const Blk &b, int *s_row, int *s_col, int np,
                                              int nt)
{
    for (int i = threadIdx.x; i < np; i += nt) {
        if (a.mode != 2 && i < b.n) {
            const size_t off = ((size_t)b.w * 2 + b.fl) * a.nstride + i;
            s_col[i] = a.part[off];
            s_row[i] = a.mode == 1 ? a.cur[off] : 0;
        } else {
            s_col[i] = 0;
            s_row[i] = 0;
        }
    }
}

// pivot key: magnitude bits (non-negative float, order-preserving) : ~row, so the
// largest key is the largest magnitude and, among equal magnitudes, the lowest row
__device__ __forceinline__ unsigned long long pivot_key(C v, int row)
{
    const float mag = (float)(fabs(v.x) + fabs(v.y));
    return ((unsigned long long)__float_as_uint(mag) << 32) | (unsigned int)(~row);
}

// Cross-lane moves without the LDS pipeline (GFX9 DPP): dpp_ctrl 0x120 + n rotates each
// 16-lane row right by n, quad_perm broadcasts one lane of each group of four.
template <int CTRL> __device__ __forceinline__ unsigned int dpp32(unsigned int v)
{
    return (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}

template <int CTRL> __device__ __forceinline__ unsigned long long dpp64(unsigned long long v)
{
    return ((unsigned long long)dpp32<CTRL>((unsigned int)(v >> 32)) << 32) | dpp32<CTRL>((unsigned int)v);
}

template <int CTRL> __device__ __forceinline__ double dppd(double v)
{
    return __longlong_as_double((long long)dpp64<CTRL>((unsigned long long)__double_as_longlong(v)));
}

__device__ __forceinline__ unsigned long long umax64(unsigned long long a, unsigned long long b) { return a > b ? a : b; }

__device__ __forceinline__ unsigned long long readlane64(unsigned long long v, int l)
{
    const unsigned int lo = (unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)v, l);
    const unsigned int hi = (unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)(v >> 32), l);
    return ((unsigned long long)hi << 32) | lo;
}

// maximum of a key over the wavefront, where the LPR lanes of a group hold the same key:
// rotations inside each 16-lane row, then the four rows' maxima through scalar registers
template <int LPR> __device__ __forceinline__ unsigned long long wave_max_key(unsigned long long key)
{
    if constexpr (LPR < 2) key = umax64(key, dpp64<0x121>(key));
    if constexpr (LPR < 4) key = umax64(key, dpp64<0x122>(key));
    if constexpr (LPR < 8) key = umax64(key, dpp64<0x124>(key));
    if constexpr (LPR < 16) key = umax64(key, dpp64<0x128>(key));
    return umax64(umax64(readlane64(key, 0), readlane64(key, 16)), umax64(readlane64(key, 32), readlane64(key, 48)));
}

// value of lane ck of each group of LPR lanes
template <int LPR, int CK> __device__ __forceinline__ C group_bcast(C v, int gbase)
{
    if constexpr (LPR == 1) {
        return v;
    } else if constexpr (LPR == 2) {
        constexpr int Q = CK | (CK << 2) | ((2 + CK) << 4) | ((2 + CK) << 6);
        return cm(dppd<Q>(v.x), dppd<Q>(v.y));
    } else if constexpr (LPR == 4) {
        constexpr int Q = CK | (CK << 2) | (CK << 4) | (CK << 6);
        return cm(dppd<Q>(v.x), dppd<Q>(v.y));
    } else {
        return cm(__shfl(v.x, gbase + CK), __shfl(v.y, gbase + CK));
    }
}

template <int LPR> __device__ __forceinline__ C group_bcast_rt(C v, int gbase, int ck)
{
    switch (ck) {
    case 0: return group_bcast<LPR, 0>(v, gbase);
    case 1: return group_bcast<LPR, 1 % LPR>(v, gbase);
    case 2: return group_bcast<LPR, 2 % LPR>(v, gbase);
    case 3: return group_bcast<LPR, 3 % LPR>(v, gbase);
    default: return cm(__shfl(v.x, gbase + ck), __shfl(v.y, gbase + ck));
    }
}

template <int FROM> __device__ __forceinline__ unsigned long long wave_max(unsigned long long key)
{
#pragma unroll
    for (int o = 32; o >= FROM; o >>= 1) {
        const unsigned long long k2 = __shfl_xor(key, o);
        key = k2 > key ? k2 : key;
    }
    return key;
}

// ln|det|, arg det and the singular flag from the pivots and the permutation, written by
// thread 0; s_red needs 3 * nw doubles.  The permutation's sign is the parity of its
// inversions (pairs i < j with perm[i] > perm[j]), counted by all threads.
template <int NT> __device__ void finish_det(const gv::DIArgs &a, const Blk &b, const C *s_piv, const int *s_perm,
                                             int np, double *s_red)
{
    constexpr int NW = NT / 64;
    double lg = 0.0, ar = 0.0;
    for (int i = threadIdx.x; i < np; i += NT) {
        const C p = s_piv[i];
        lg += log(hypot(p.x, p.y));
        ar += atan2(p.y, p.x);
    }
    unsigned int inv = 0;
    for (int e = threadIdx.x; e < np * np; e += NT) {
        const int i = e / np, j = e - i * np;
        inv += (unsigned int)(j > i && s_perm[j] < s_perm[i]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lg += __shfl_xor(lg, o);
        ar += __shfl_xor(ar, o);
    }
    const unsigned int odd = (unsigned int)__popcll(__ballot(inv & 1u)) & 1u;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        s_red[wid] = lg;
        s_red[NW + wid] = ar;
        s_red[2 * NW + wid] = (double)odd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double la = 0.0, ph = 0.0;
        int odd_all = 0;
        for (int q = 0; q < NW; ++q) {
            la += s_red[q];
            ph += s_red[NW + q];
            odd_all ^= (int)s_red[2 * NW + q];
        }
        if (odd_all) ph += M_PI;
        a.logabs[b.o] = la;
        a.phase[b.o] = isfinite(ph) ? remainder(ph, 2 * M_PI) : 0.0;
        a.status[b.o] = !(2.0 * la >= -20.72326583694641);  // |det|^2 < 1e-9 (ln 1e-9), NaN included
    }
}

// the largest divisor of n up to m: slots per batch of the update
constexpr int batch_slots(int n, int m)
{
    for (int d = m; d > 1; --d)
        if (n % d == 0) return d;
    return 1;
}

template <int NP, int LPR>
__global__ void __launch_bounds__(((NP * LPR + 63) / 64) * 64) di_reg_kernel(gv::DIArgs a)
{
    constexpr int NT = ((NP * LPR + 63) / 64) * 64;
    constexpr int NW = NT / 64;
    constexpr int CPT = NP / LPR;
    constexpr int BT = batch_slots(CPT, CPT >= 32 ? 4 : 8);
    static_assert(NP % LPR == 0 && (LPR & (LPR - 1)) == 0, "LPR: a power of two dividing NP");

    __shared__ C s_prow[2][NP];      // the pivot row, in (slot, lane) order t * LPR + c
    __shared__ unsigned long long s_ckey[2][NW];
    __shared__ C s_cpiv[2][NW];      // each wave's candidate pivot value
    __shared__ C s_piv[NP];
    __shared__ int s_perm[NP];       // pivot row of column k
    __shared__ int s_invp[NP];       // column at which row i was pivoted
    __shared__ int s_row[NP], s_col[NP];
    __shared__ double s_red[3 * NW];

    const Blk b = blk_of(a);
    const int n = b.n;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int row = tid / LPR, c = tid % LPR;
    const bool has_row = row < NP;
    stage_indices(a, b, s_row, s_col, NP, NT);
    __syncthreads();

    // thread (row, c) holds column c + LPR t in slot t; rows and columns past n are the
    // identity
    C u[CPT];
#pragma unroll
    for (int t = 0; t < CPT; ++t) {
        const int j = c + LPR * t;
        u[t] = cm(row == j ? 1.0 : 0.0, 0.0);
        if (row < n && j < n) u[t] = elem(a, b, s_row, s_col, row, j);
    }
    bool pivoted = !has_row;
    C sc = cm(1.0, 0.0);
    int par = 0;
    const int gbase = tid & ~(LPR - 1);
    // After every LPR columns each row's slots rotate left by one, so the column being
    // eliminated is always slot 0 of lane ck (a compile-time register index).
    for (int kk = 0; kk < CPT; ++kk) {
#pragma unroll 1
        for (int ck = 0; ck < LPR; ++ck) {
            const int k = ck + LPR * kk;
            const C uk = group_bcast_rt<LPR>(u[0], gbase, ck);
            unsigned long long key = pivoted ? 0ull : pivot_key(uk, row);
            key = wave_max_key<LPR>(key);  // uniform over the wave
            // Only the keys and candidate pivot values cross the first barrier; the winning
            // row alone is written after it.  (A ds_write costs the LDS pipeline the whole
            // wave's slots however few lanes are active, so every wave writing its candidate
            // row would cost 8x the LDS time of one row.)
            const int wrow = (int)(~(unsigned int)(key & 0xffffffffull));
            if (!pivoted && row == wrow && c == 0) s_cpiv[par][wid] = uk;
            if (lane == 0) s_ckey[par][wid] = key;
            __syncthreads();
            unsigned long long bkey = s_ckey[par][0];
            C piv = s_cpiv[par][0];
#pragma unroll
            for (int q = 1; q < NW; ++q) {
                const unsigned long long v = s_ckey[par][q];
                const C pq = s_cpiv[par][q];
                if (v > bkey) {
                    bkey = v;
                    piv = pq;
                }
            }
            // NP - k >= 1 rows are unpivoted, so bkey names a row
            const int p = (int)(~(unsigned int)(bkey & 0xffffffffull));
            if (row == p) {
#pragma unroll
                for (int t = 0; t < CPT; ++t) s_prow[par][t * LPR + c] = u[t];
            }
            const C *pr = &s_prow[par][c];
            const double np2 = fma(piv.x, piv.x, piv.y * piv.y);
            const double rn = np2 > 0.0 ? 1.0 / np2 : 0.0;
            const C ainv = cm(piv.x * rn, -piv.y * rn);
            if (tid == 0) {
                s_perm[k] = p;
                s_invp[p] = k;
                s_piv[k] = piv;
            }
            const bool me = row == p;
            C f = cmul(uk, ainv);
            if (me) {
                f = cm(0.0, 0.0);
                sc = cm(-ainv.x, -ainv.y);
                pivoted = true;
            }
            // the multiplier (and its division) is formed while the pivot row is written;
            // the second barrier publishes the row
            asm volatile("" ::"v"(f.x), "v"(f.y));
            __syncthreads();
            // U -= f r in batches of BT slots, the next batch's LDS reads in flight (the
            // compiler barrier keeps the reads from all being hoisted, which would need
            // the whole pivot row in VGPRs next to U)
            C r0[BT], r1[BT];
#pragma unroll
            for (int q = 0; q < BT; ++q) r0[q] = pr[q * LPR];
#pragma unroll
            for (int tb = 0; tb < CPT; tb += BT) {
                if (tb + BT < CPT) {
#pragma unroll
                    for (int q = 0; q < BT; ++q) r1[q] = pr[(tb + BT + q) * LPR];
                }
                asm volatile("" ::: "memory");
#pragma unroll
                for (int q = 0; q < BT; ++q) u[tb + q] = cfms(f, r0[q], u[tb + q]);
#pragma unroll
                for (int q = 0; q < BT; ++q) r0[q] = r1[q];
            }
            if (c == ck) u[0] = me ? cm(-1.0, 0.0) : f;
            par ^= 1;
        }
        const C first = u[0];
#pragma unroll
        for (int t = 0; t + 1 < CPT; ++t) u[t] = u[t + 1];
        u[CPT - 1] = first;
    }
    __syncthreads();
    // inverse: (A^-1)[invp[i]][perm[j]] = s_i U[i][j]; the slots are back in their
    // original order after CPT rotations.  Written before the determinant's logarithms,
    // which would otherwise share the register file with U.
    if (row < n) {
        const int x = s_invp[row];
#pragma unroll
        for (int t = 0; t < CPT; ++t) {
            const int j = c + LPR * t;
            if (j < n) {
                const int y = s_perm[j];
                if (x < n && y < n) put(a, b, x, y, cmul(sc, u[t]));
            }
        }
    }
    finish_det<NT>(a, b, s_piv, s_perm, NP, s_red);
}

// a double of lane l (l uniform), into scalar registers
__device__ __forceinline__ double rdl(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const unsigned int lo = (unsigned int)__builtin_amdgcn_readlane((int)b, l);
    const unsigned int hi = (unsigned int)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Panel-blocked form of di_reg_kernel's elimination (same pivots, same multipliers and the
// same multiply-adds in the same order per element, so the same bits).  Wave w holds the
// PW columns [PW w, PW w + PW) of all rows, lane l the rows l + 64 m.  Columns are eliminated
// in panels of PW, panel P by wave P alone: its pivot search is a wave reduction, the pivot
// row's values reach its lanes through scalar registers (readlane), and each step's
// multipliers go to LDS.  Every other wave applies the panel's steps to its own columns in
// order: the pivot row of step k is row perm[k] of its own registers (already updated
// through step k - 1), again broadcast by readlane, and the multipliers come from LDS.
//
// No barriers between the first and the last column: the producing wave publishes each
// step with a release store of the step count, and the other waves follow step by step
// with acquire loads, so that wave P + 1 (raised issue priority) starts factoring panel
// P + 1 one step after panel P's last step.  The multipliers live in a ring of NB panels;
// a producer first waits until every wave is done with the panel it overwrites.  Every
// wait is bounded: past kSpin polls (about 4 ms; the longest wait in a correct run is a few
// panels' work, under 0.1 ms) the wave stops waiting for the rest of the matrix and the
// matrix is flagged singular (status 1), so that no logic error can keep a workgroup from
// finishing.
constexpr int kSpin = 1 << 16;

template <int NP, int PW>
__global__ void __launch_bounds__(NP / PW * 64) di_pan_kernel(gv::DIArgs a)
{
    constexpr int NW = NP / PW;
    constexpr int NT = NW * 64;
    constexpr int RPL = (NP + 63) / 64;
    constexpr int NB = 4;
    static_assert(NP % PW == 0 && NT <= 1024 && NW <= 64, "NP: a multiple of PW, at most 16 waves");

    __shared__ C s_F[NB][PW][RPL * 64];  // multipliers of the ring's panels, by row
    __shared__ C s_piv[NP];
    __shared__ C s_sc[RPL * 64];         // -1/pivot of each pivot row
    __shared__ int s_perm[NP];           // pivot row of column k
    __shared__ int s_invp[RPL * 64];     // column at which row i was pivoted
    __shared__ int s_done[RPL * 64];     // row pivoted (rows past NP count as pivoted)
    __shared__ int s_row[NP], s_col[NP];
    __shared__ int s_sync[2 + NW];       // steps published, wait failure, panels done per wave
    __shared__ double s_red[3 * NW];

    const Blk b = blk_of(a);
    const int n = b.n;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    stage_indices(a, b, s_row, s_col, NP, NT);
    for (int i = tid; i < RPL * 64; i += NT) {
        s_done[i] = i >= NP;
        s_sc[i] = cm(1.0, 0.0);
        s_invp[i] = 0;
    }
    for (int i = tid; i < 2 + NW; i += NT) s_sync[i] = 0;
    __syncthreads();

    // rows and columns past n are the identity, rows past NP zero
    C u[RPL][PW];
#pragma unroll
    for (int m = 0; m < RPL; ++m) {
        const int row = m * 64 + lane;
#pragma unroll
        for (int j = 0; j < PW; ++j) {
            const int col = wv * PW + j;
            u[m][j] = cm(row == col ? 1.0 : 0.0, 0.0);
            if (row < n && col < n) u[m][j] = elem(a, b, s_row, s_col, row, col);
        }
    }
    auto ld_acq = [&](int i) __attribute__((always_inline)) {
        return __hip_atomic_load(&s_sync[i], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    auto st_rel = [&](int i, int v) __attribute__((always_inline)) {
        __hip_atomic_store(&s_sync[i], v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    bool failed = false;
    int seen = 0;  // steps known to be published
    auto wait_pub = [&](int need) __attribute__((always_inline)) {
        if (seen >= need || failed) return;
        for (int it = 0; it < kSpin; ++it) {
            seen = __builtin_amdgcn_readfirstlane(ld_acq(0));
            if (seen >= need) return;
            __builtin_amdgcn_s_sleep(1);
        }
        failed = true;
        seen = need;
    };
    // before panel P's multipliers overwrite ring slot P % NB: every wave done with P - NB
    auto wait_ring = [&](int P) __attribute__((always_inline)) {
        const int need = P - NB + 1;
        if (need <= 0 || failed) return;
        for (int it = 0; it < kSpin; ++it) {
            const int v = lane < NW ? ld_acq(2 + lane) : need;
            if (__ballot(v < need) == 0ull) return;
            __builtin_amdgcn_s_sleep(1);
        }
        failed = true;
    };
    // the pivot row's values of this wave's columns: scalar registers
    auto prow = [&](int p, C (&r)[PW]) __attribute__((always_inline)) {
        const int pl = p & 63, pm = p >> 6;
#pragma unroll
        for (int m = 0; m < RPL; ++m)
            if (m == pm) {
#pragma unroll
                for (int j = 0; j < PW; ++j) r[j] = cm(rdl(u[m][j].x, pl), rdl(u[m][j].y, pl));
            }
    };
    // factorisation of panel P by wave P: its PW steps on its own columns, each published
    auto factor = [&](int P) __attribute__((always_inline)) {
        const int slot = P % NB;
        bool pv[RPL];
#pragma unroll
        for (int m = 0; m < RPL; ++m) pv[m] = s_done[m * 64 + lane] != 0;
#pragma unroll
        for (int q = 0; q < PW; ++q) {
            const int k = P * PW + q;
            unsigned long long key = 0ull;
#pragma unroll
            for (int m = 0; m < RPL; ++m)
                if (!pv[m]) key = umax64(key, pivot_key(u[m][q], m * 64 + lane));
            key = wave_max_key<1>(key);
            // NP - k >= 1 rows are unpivoted, so the key names a row
            const int p = (int)(~(unsigned int)(key & 0xffffffffull));
            C r[PW];
            prow(p, r);
            const C piv = r[q];
            const double np2 = fma(piv.x, piv.x, piv.y * piv.y);
            const double rn = np2 > 0.0 ? 1.0 / np2 : 0.0;
            const C ainv = cm(piv.x * rn, -piv.y * rn);
            C f[RPL];
            bool me[RPL];
#pragma unroll
            for (int m = 0; m < RPL; ++m) {
                me[m] = m * 64 + lane == p;
                f[m] = me[m] ? cm(0.0, 0.0) : cmul(u[m][q], ainv);
                pv[m] = pv[m] || me[m];
                s_F[slot][q][m * 64 + lane] = f[m];
            }
            if (lane == 0) {
                s_perm[k] = p;
                s_invp[p] = k;
                s_piv[k] = piv;
                s_sc[p] = cm(-ainv.x, -ainv.y);
                s_done[p] = 1;
                st_rel(0, k + 1);
            }
            // the next step's pivot column first: its pivot search can then start while the
            // other columns are updated
            if (q + 1 < PW) {
#pragma unroll
                for (int m = 0; m < RPL; ++m) u[m][q + 1] = cfms(f[m], r[q + 1], u[m][q + 1]);
            }
#pragma unroll
            for (int m = 0; m < RPL; ++m) {
#pragma unroll
                for (int j = 0; j < PW; ++j)
                    if (j != q + 1) u[m][j] = cfms(f[m], r[j], u[m][j]);
                u[m][q] = me[m] ? cm(-1.0, 0.0) : f[m];
            }
        }
    };
    // panel P's PW steps on this wave's columns, in order, each once it is published
    auto apply = [&](int P) __attribute__((always_inline)) {
        const int slot = P % NB;
#pragma unroll 1
        for (int q = 0; q < PW; ++q) {
            const int k = P * PW + q;
            wait_pub(k + 1);
            const int p = __builtin_amdgcn_readfirstlane(s_perm[k]);
            C f[RPL];
#pragma unroll
            for (int m = 0; m < RPL; ++m) f[m] = s_F[slot][q][m * 64 + lane];
            C r[PW];
            prow(p, r);
#pragma unroll
            for (int m = 0; m < RPL; ++m)
#pragma unroll
                for (int j = 0; j < PW; ++j) u[m][j] = cfms(f[m], r[j], u[m][j]);
        }
    };
#pragma unroll 1
    for (int P = 0; P < NW; ++P) {
        if (wv == P + 1) __builtin_amdgcn_s_setprio(2);
        if (wv != P) apply(P);
        if (wv == P) {
            wait_ring(P);
            factor(P);
            __builtin_amdgcn_s_setprio(0);
        }
        if (lane == 0) st_rel(2 + wv, P + 1);
    }
    if (failed && lane == 0) s_sync[1] = 1;
    __syncthreads();
    // inverse: (A^-1)[invp[i]][perm[j]] = sc_i U[i][j]
#pragma unroll
    for (int m = 0; m < RPL; ++m) {
        const int row = m * 64 + lane;
        if (row < n) {
            const int x = s_invp[row];
            const C sc = s_sc[row];
#pragma unroll
            for (int j = 0; j < PW; ++j) {
                const int col = wv * PW + j;
                if (col < n) {
                    const int y = s_perm[col];
                    if (x < n && y < n) put(a, b, x, y, cmul(sc, u[m][j]));
                }
            }
        }
    }
    finish_det<NT>(a, b, s_piv, s_perm, NP, s_red);
    if (tid == 0 && s_sync[1]) a.status[b.o] = 1;
}

// One row per thread (n <= 256), panels of NB columns; work: column-major, leading
// dimension a.ld, element (row, col) at work[block][col * ld + row].
template <int NB>
__global__ void __launch_bounds__(256) di_stream_kernel(gv::DIArgs a)
{
    constexpr int NT = 256, NW = 4;
    __shared__ C s_R[NB][NT];              // pivot rows of the panel, value at each column
    __shared__ C s_cand[2][NW][2 * NB];    // each wave's candidate: panel values, then multipliers
    __shared__ unsigned long long s_ckey[2][NW];
    __shared__ C s_piv[NT];
    __shared__ int s_perm[NT], s_invp[NT];
    __shared__ int s_row[NT], s_col[NT];
    __shared__ double s_red[3 * NW];

    const Blk b = blk_of(a);
    const int n = b.n;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int row = tid;
    const bool has_row = row < n;
    const size_t ld = (size_t)a.ld;
    C *Wm = a.work + (size_t)blockIdx.x * ld * (size_t)n;
    stage_indices(a, b, s_row, s_col, NT, NT);
    __syncthreads();
    if (has_row)
        for (int j = 0; j < n; ++j) Wm[(size_t)j * ld + row] = elem(a, b, s_row, s_col, row, j);
    // each thread reads back only what it wrote until the first barrier of a panel
    bool pivoted = !has_row;
    C sc = cm(1.0, 0.0);
    int par = 0;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = n - k0 < NB ? n - k0 : NB;
        C P[NB], F[NB];
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            P[m] = (has_row && m < nb) ? Wm[(size_t)(k0 + m) * ld + row] : cm(0.0, 0.0);
            F[m] = cm(0.0, 0.0);
        }
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            if (m < nb) {
                const int k = k0 + m;
                unsigned long long key = pivoted ? 0ull : pivot_key(P[m], row);
                key = wave_max<1>(key);
                const int wrow = (int)(~(unsigned int)(key & 0xffffffffull));
                if (!pivoted && row == wrow) {
#pragma unroll
                    for (int q = 0; q < NB; ++q) {
                        s_cand[par][wid][q] = P[q];
                        s_cand[par][wid][NB + q] = F[q];
                    }
                }
                if (lane == 0) s_ckey[par][wid] = key;
                __syncthreads();
                int bw = 0;
                unsigned long long bkey = s_ckey[par][0];
#pragma unroll
                for (int q = 1; q < NW; ++q) {
                    const unsigned long long v = s_ckey[par][q];
                    if (v > bkey) {
                        bkey = v;
                        bw = q;
                    }
                }
                const int p = (int)(~(unsigned int)(bkey & 0xffffffffull));
                const C *cand = s_cand[par][bw];
                const C piv = cand[m];
                const double np2 = fma(piv.x, piv.x, piv.y * piv.y);
                const C ainv = np2 > 0.0 ? cm(piv.x / np2, -piv.y / np2) : cm(0.0, 0.0);
                if (tid == 0) {
                    s_perm[k] = p;
                    s_invp[p] = k;
                    s_piv[k] = piv;
                }
                // the pivot row at this step, outside the panel: its panel-start value
                // minus the earlier panel steps (thread t rebuilds column t)
                if (tid < n && (tid < k0 || tid >= k0 + nb)) {
                    C r = Wm[(size_t)tid * ld + p];
#pragma unroll
                    for (int q = 0; q < NB; ++q)
                        if (q < m) r = cfms(cand[NB + q], s_R[q][tid], r);
                    s_R[m][tid] = r;
                }
                const bool me = row == p;
                C f = cmul(P[m], ainv);
                if (me) {
                    f = cm(0.0, 0.0);
                    sc = cm(-ainv.x, -ainv.y);
                    pivoted = true;
                }
#pragma unroll
                for (int q = 0; q < NB; ++q) P[q] = cfms(f, cand[q], P[q]);
                P[m] = me ? cm(-1.0, 0.0) : f;
                F[m] = f;
                par ^= 1;
            }
        }
        __syncthreads();
        // the rest of the row: rank-nb update with the panel's pivot rows
        if (has_row) {
            for (int j = 0; j < n; ++j) {
                if (j >= k0 && j < k0 + nb) continue;
                C v = Wm[(size_t)j * ld + row];
#pragma unroll
                for (int q = 0; q < NB; ++q)
                    if (q < nb) v = cfms(F[q], s_R[q][j], v);
                Wm[(size_t)j * ld + row] = v;
            }
#pragma unroll
            for (int m = 0; m < NB; ++m)
                if (m < nb) Wm[(size_t)(k0 + m) * ld + row] = P[m];
        }
        // the next panel's rebuild reads other rows of the work buffer and overwrites s_R
        __threadfence_block();
        __syncthreads();
    }
    finish_det<NT>(a, b, s_piv, s_perm, n, s_red);
    if (has_row) {
        const int x = s_invp[row];
        for (int j = 0; j < n; ++j) {
            const int y = s_perm[j];
            if (x < n && y < n) put(a, b, x, y, cmul(sc, Wm[(size_t)j * ld + row]));
        }
    }
}

template <int NP, int LPR> static hipError_t launch_reg(const gv::DIArgs &a, int nblocks, hipStream_t st)
{
    constexpr int NT = ((NP * LPR + 63) / 64) * 64;
    hipLaunchKernelGGL((di_reg_kernel<NP, LPR>), dim3(nblocks), dim3(NT), 0, st, a);
    return hipGetLastError();
}

template <int NP, int PW> static hipError_t launch_pan(const gv::DIArgs &a, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL((di_pan_kernel<NP, PW>), dim3(nblocks), dim3(NP / PW * 64), 0, st, a);
    return hipGetLastError();
}

struct RegShape {
    int np;
    hipError_t (*fn)(const gv::DIArgs &, int, hipStream_t);
};

// padded sizes: the ground-state N of L = 4, 6, 8, 10, 12, 16 exactly, and the channel
// sizes L^2/2 + 1 of the transverse exciton at L <= 12 (9, 19, 33, 51, 73) rounded up;
// larger matrices (n > 128) take the streamed kernel
static const RegShape kReg[] = {
    {8, launch_reg<8, 8>},    {12, launch_reg<12, 4>},  {18, launch_reg<18, 2>},  {20, launch_reg<20, 4>},
    {32, launch_reg<32, 4>},  {36, launch_reg<36, 4>},  {50, launch_reg<50, 2>},  {52, launch_reg<52, 4>},
    {72, launch_reg<72, 4>},  {76, launch_reg<76, 4>},  {128, launch_pan<128, 8>},
};

constexpr int kStreamNB = 8;
constexpr int kStreamMax = 256;

}  // namespace gd

extern "C" int gv_detinv_supported(int n) { return n >= 1 && n <= gd::kStreamMax; }

extern "C" size_t gv_detinv_work_bytes(int n)
{
    for (const auto &e : gd::kReg)
        if (n <= e.np) return 0;
    const size_t ld = ((size_t)n + 63) / 64 * 64;
    return ld * (size_t)n * sizeof(double2);
}

extern "C" int gv_detinv_ld(int n) { return (n + 63) / 64 * 64; }

extern "C" hipError_t gv_launch_detinv(const gv::DIArgs *a, int nblocks, hipStream_t st)
{
    const int n = a->mode == 2 ? a->n_up : (a->n_up > a->n_dn ? a->n_up : a->n_dn);
    if (n < 1 || n > gd::kStreamMax || nblocks < 1) return hipErrorInvalidValue;
    for (const auto &e : gd::kReg)
        if (n <= e.np) return e.fn(*a, nblocks, st);
    if (!a->work || a->ld < n) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gd::di_stream_kernel<gd::kStreamNB>, dim3(nblocks), dim3(256), 0, st, *a);
    return hipGetLastError();
}

// ---- C ABI: batched linalg::DetInv on caller device memory -----------------------------

namespace {

constexpr int kMaxDevices = 64;

// per-device workspace of gpvmc_det_inv (the streamed kernel's working matrices), kept
// and grown across calls
struct Workspace {
    void *p = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mu;
Workspace g_ws[kMaxDevices];

int detinv_check(int32_t device, int32_t n, int32_t batch, const double *a, double *ai, double *logabs,
                 double *phase, int32_t *singular)
{
    if (!a || !ai || !logabs || !phase || !singular) return gv_fail(GPVMC_ERR_ARG, "null argument");
    if (n < 1 || !gv_detinv_supported(n))
        return gv_fail(GPVMC_ERR_UNSUPPORTED, ("matrix size " + std::to_string(n) + " outside 1 .. 256").c_str());
    if (batch < 1) return gv_fail(GPVMC_ERR_ARG, "batch must be >= 1");
    if (device < 0 || device >= kMaxDevices) return gv_fail(GPVMC_ERR_ARG, "device ordinal out of range");
    return GPVMC_OK;
}

int detinv_launch(int32_t n, int32_t batch, const double *a, double *ai, double *logabs, double *phase,
                  int32_t *singular, void *work, void *stream)
{
    gv::DIArgs d{};
    d.mode = 2;
    d.n_up = n;
    d.n_dn = n;
    d.ld = gv_detinv_ld(n);
    d.A = reinterpret_cast<const double2 *>(a);
    d.Ai64 = reinterpret_cast<double2 *>(ai);
    d.logabs = logabs;
    d.phase = phase;
    d.status = singular;
    d.work = static_cast<double2 *>(work);
    const hipError_t e = gv_launch_detinv(&d, batch, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return gv_fail(GPVMC_ERR_HIP, (std::string("det_inv launch: ") + hipGetErrorString(e)).c_str());
    return GPVMC_OK;
}

}  // namespace

extern "C" int gpvmc_det_inv(int32_t device, int32_t n, int32_t batch, const double *a, double *ai,
                             double *logabs, double *phase, int32_t *singular, void *stream)
{
    int rc = detinv_check(device, n, batch, a, ai, logabs, phase, singular);
    if (rc) return rc;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return gv_fail(GPVMC_ERR_HIP, (std::string("hipSetDevice: ") + hipGetErrorString(e)).c_str());
    void *work = nullptr;
    if (const size_t wb = gv_detinv_work_bytes(n) * (size_t)batch) {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        Workspace &ws = g_ws[device];
        if (ws.bytes < wb) {
            if (ws.p) (void)hipFree(ws.p);
            ws.p = nullptr;
            ws.bytes = 0;
            e = hipMalloc(&ws.p, wb);
            if (e != hipSuccess)
                return gv_fail(GPVMC_ERR_HIP, (std::string("det_inv workspace: ") + hipGetErrorString(e)).c_str());
            ws.bytes = wb;
        }
        work = ws.p;
    }
    return detinv_launch(n, batch, a, ai, logabs, phase, singular, work, stream);
}

extern "C" int gpvmc_det_inv_workspace_bytes(int32_t n, int64_t *bytes)
{
    if (!bytes) return gv_fail(GPVMC_ERR_ARG, "null argument");
    if (n < 1 || !gv_detinv_supported(n))
        return gv_fail(GPVMC_ERR_UNSUPPORTED, ("matrix size " + std::to_string(n) + " outside 1 .. 256").c_str());
    *bytes = (int64_t)gv_detinv_work_bytes(n);
    return GPVMC_OK;
}

// ---- allocation-free variant ------------------------------------------------------------
// For callers that cannot allocate device memory (code under graph capture, serving loops):
// the streamed kernel's workspace is a fixed file-scope device array.  Each device holds
// its own copy until the process ends; its address is looked up once per device and used
// in place of the array's host-side name.

namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;  // 64 MiB per device

__device__ __align__(256) char scratch[kNoAllocScratchBytes];

std::mutex g_scr_mu;
void *g_scr_ptr[kMaxDevices];  // device address of `scratch` on each device

// the device address of `scratch` on `device` (current), looked up on the first call
int scratch_ptr(int32_t device, void **out)
{
    std::lock_guard<std::mutex> lk(g_scr_mu);
    if (!g_scr_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return gv_fail(GPVMC_ERR_HIP, (std::string("scratch address: ") + hipGetErrorString(e)).c_str());
        g_scr_ptr[device] = p;
    }
    *out = g_scr_ptr[device];
    return GPVMC_OK;
}

}  // namespace

extern "C" int gpvmc_det_inv_noalloc(int32_t device, int32_t n, int32_t batch, const double *a, double *ai,
                                     double *logabs, double *phase, int32_t *singular, void *stream)
{
    int rc = detinv_check(device, n, batch, a, ai, logabs, phase, singular);
    if (rc) return rc;
    const size_t per = gv_detinv_work_bytes(n);
    if (per * (size_t)batch > kNoAllocScratchBytes)
        return gv_fail(GPVMC_ERR_ARG,
                       ("gpvmc_det_inv_noalloc: " + std::to_string(batch) + " matrices of size " + std::to_string(n) +
                        " need " + std::to_string(per * (size_t)batch) + " bytes of workspace, above its fixed " +
                        std::to_string(kNoAllocScratchBytes) + "; at most " +
                        std::to_string(kNoAllocScratchBytes / per) + " per call at this size")
                           .c_str());
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return gv_fail(GPVMC_ERR_HIP, (std::string("hipSetDevice: ") + hipGetErrorString(e)).c_str());
    void *work = nullptr;
    if (per) {
        rc = scratch_ptr(device, &work);
        if (rc) return rc;
    }
    return detinv_launch(n, batch, a, ai, logabs, phase, singular, work, stream);
}
