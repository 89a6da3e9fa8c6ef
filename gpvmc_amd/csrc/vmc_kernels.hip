// gfx950 kernels of the Metropolis VMC hot path: one workgroup per Markov chain.
//
// Thread layout (shared by the sweep and the Gauss-Jordan kernels): thread t owns row
// `row = t / LPR` of each flavour's N x N matrix and the LPR-strided columns
// j = c + LPR*k (c = t % LPR, k < CPT = N/LPR).  A row dot product is CPT complex FMAs per
// lane plus a log2(LPR)-stage xor-shuffle reduction inside the row's lane group, so the
// ratio of a proposal (one row of the inverse against one orbital column) costs O(N/LPR)
// on the critical path, and the Sherman-Morrison update is one fused read-modify-write
// pass over the inverse with no cross-row communication except one LDS broadcast of the
// replaced row.  Workgroups are padded to whole wavefronts; padding threads own no row.
//
// Matrix conventions (reference: SlaterDeterminant.cpp:68-74, linalg.cpp:9-44):
// A(f, r) = Phi(orbital of wavefunction particle f, site of lattice particle r); the
// inverse is stored row-major as Ai[label r][slot f] = (A^-1)(r, f); the ratio of moving
// lattice particle ci to site s is q = sum_f Ai[ci][f] Phi(orb_f, s)  (linalg.cpp:143-154).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vmc_kernels.h"

namespace gv {

template <typename T> struct cpx;
template <> struct cpx<float> { using t = float2; };
template <> struct cpx<double> { using t = double2; };

template <typename C> __device__ __forceinline__ C cmk(double x, double y)
{
    C r;
    r.x = x;
    r.y = y;
    return r;
}

template <typename C> __device__ __forceinline__ C cfma(C a, C b, C acc)
{
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
    return acc;
}

// complex product with a fixed FMA order (the same rounding wherever it is inlined)
template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    C r;
    r.x = fma(a.x, b.x, -(a.y * b.y));
    r.y = fma(a.x, b.y, a.y * b.x);
    return r;
}

// sum over the LPR lanes of one row group (aligned, power-of-two group inside a wavefront)
template <int LPR, typename C> __device__ __forceinline__ C row_reduce(C v)
{
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) {
        v.x += __shfl_xor(v.x, o);
        v.y += __shfl_xor(v.y, o);
    }
    return v;
}

__device__ __forceinline__ double u32_to_open01(uint32_t x)
{
    return ((double)x + 0.5) * 2.3283064365386963e-10;  // 2^-32
}

// block-wide sum of NV doubles; every thread receives the totals.
template <int NT, int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*s_red)[4])
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[i] += __shfl_xor(v[i], o);
    }
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) s_red[wid][i] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < NW; ++q) t += s_red[q][i];
        v[i] = t;
    }
}

// one DPP move of a double (two 32-bit halves); lanes of rows outside ROW_MASK read 0
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_f64(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo));
}

// sum over the 64 lanes of a wavefront, left in lane 63 (every lane active): quad
// permutations, half-row and row mirrors, then the row broadcasts 15 and 31
__device__ __forceinline__ double wave_sum_to_lane63(double v)
{
    v += dpp_f64<0xB1>(v);        // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4E>(v);        // quad_perm [2, 3, 0, 1]
    v += dpp_f64<0x141>(v);       // row_half_mirror
    v += dpp_f64<0x140>(v);       // row_mirror: every lane holds its row's sum
    v += dpp_f64<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

constexpr int isqrt_c(int n)
{
    int r = 0;
    while ((r + 1) * (r + 1) <= n) ++r;
    return r;
}

// ---------------------------------------------------------------------------------
// Wave-uniform helpers of the sweep kernel.

// Register vectors: element k of a vector is register (base + k), so an element picked
// with a wave-uniform index compiles to one indexed move (s_set_gpr_idx_on + v_mov)
// and the vector never leaves registers.
template <typename T, int M> using regvec = T __attribute__((ext_vector_type(M)));

// largest divisor d of n with d * unit <= budget (at least 1)
constexpr int divisor_within(int n, int unit, int budget)
{
    int best = 1;
    for (int d = 1; d <= n; ++d)
        if (n % d == 0 && d * unit <= budget) best = d;
    return best;
}

// v_readlane of a float or double (wave-uniform lane index)
__device__ __forceinline__ float readlane_real(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane_real(double v, int l)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// amplitude mantissa (mr, mi) *= q.  The mantissa is rescaled by a power of two (exact) only
// when its binary exponent leaves [-256, 256]; the exponent removed is added to *ex, so
// amplitude = (mr + i mi) 2^ex throughout (the reference's BigComplex, BigComplex.h)
__device__ __forceinline__ void amp_mul(double &mr, double &mi, double &ex, double qr, double qi)
{
    const double re = mr * qr - mi * qi, im = mr * qi + mi * qr;
    const int be = (int)((__double_as_longlong(fmax(fabs(re), fabs(im))) >> 52) & 0x7FF) - 1023;
    if (be > 256 || be < -256) {
        const double sc = __longlong_as_double((long long)(1023 - be) << 52);  // 2^-be
        mr = re * sc;
        mi = im * sc;
        ex += be;
    } else {
        mr = re;
        mi = im;
    }
}

// number of set bits of the label mask m below bit x (x < 64 * NMW)
template <int NMW> __device__ __forceinline__ int mask_rank(const uint64_t (&m)[NMW], int x)
{
    if constexpr (NMW == 1) {
        return __popcll(m[0] & ((1ull << x) - 1ull));
    } else {
        return x < 64 ? __popcll(m[0] & ((1ull << x) - 1ull))
                      : __popcll(m[0]) + __popcll(m[1] & ((1ull << (x - 64)) - 1ull));
    }
}

// orders this wavefront's LDS accesses (one wavefront's LDS operations complete in issue
// order; this keeps the compiler from moving them across)
__device__ __forceinline__ void wave_lds_order()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Delayed-update windows (WDU sweep kernel), sizes of the dynamic LDS carve-up:
//   two row dumps [WD_ROWS][LPR][2][RP] of T (double-buffered by window parity),
//   then the window scratch of wavefront 0 (WduScratch).
constexpr int WD_K = 16;     // steps per window
constexpr int WD_M = 4;      // accepted moves per window at most
constexpr int WD_ROWS = 2 * WD_K;
template <typename T> struct WduScratch {
    using C = typename cpx<T>::t;
    // vectors of move n by sub-block position: [0] x_up conj(q)/|q|^2 over U, [1] y_up over
    // D, [2] x_dn conj(q)/|q|^2 over D, [3] y_dn over U (the rank-1 factors of the move on
    // each table, the column one times 1/q as the table update forms it)
    C vec[WD_M][4][WD_ROWS];
    C xr[WD_M][2][WD_ROWS];           // the column factors x_up, x_dn themselves
    double2 iq[2][WD_M];              // 1 / q in FP64, for the coefficients
    double2 qd[2][WD_M];              // q in FP64, for the amplitude
    double2 coef[2][2][WD_M][WD_M];   // [table][gamma, beta][n][m]
    double2 st[2][2][WD_M][WD_M];     // [table][S, T][n][d]
    C mm[2][WD_M][WD_M];              // M = S^T diag(1/q) T per table
    int lab[2][WD_ROWS];              // labels of the window by position, per flavour
    int mv[WD_M][4];                  // ci, cj, sa, sb of each move
    int pos[WD_M][2];                 // position of ci in U, of cj in D
    int res[4];                       // moves, steps processed
    uint32_t msk[16][2][4];           // per wavefront: label masks of the window's sites
};
template <typename T, int N, int LPR> constexpr size_t wdu_dump_elems()
{
    return (size_t)WD_ROWS * LPR * 2 * (N / LPR + 16 / sizeof(T));
}
template <typename T, int N, int LPR> constexpr size_t wdu_lds_bytes()
{
    return 2 * wdu_dump_elems<T, N, LPR>() * sizeof(T) + sizeof(WduScratch<T>);
}

// ---------------------------------------------------------------------------------
// Sweep kernel: n_steps of MetroMC::Step (MetroMC.cpp:90-103) for one walker per block.
//
// State: the ratio tables G_fl[i][j] = sum_f Ai_fl[i][f] Phi_fl(orb_f, site of particle j
// of the other flavour) = det(A_fl with particle i moved to that site) / det(A_fl).  For
// the swap of up particle ci with down particle cj (LatticeStepper.cpp:70-73) the two
// single-hop ratios of linalg::DetUpdate (linalg.cpp:143-154) are table entries,
// q_up = G_up[ci][cj] and q_dn = G_dn[cj][ci]; the Sherman-Morrison update of the inverse
// (linalg.cpp:402-561) becomes the rank-1 update G -= (G[:,cj] - e_ci)(G[ci,:] + e_cj)/q
// of the table (column cj then describes the site the moved-away particle left).  The
// ProjHeis swap ratios are table products G_up[i][j] G_dn[j][i].
// RES = true keeps both tables in registers for the whole launch; RES = false reads and
// writes the walker's rows in global memory at every use.
//
// Windows.  A rejected or trivial step leaves configuration and tables unchanged, so the
// proposals of the following steps can all be evaluated against the current tables.  The
// kernel therefore advances by windows of up to KW steps:
//  * proposals come from a per-wave batch: every 64 steps lane l draws the uniforms of
//    step s + l (Philox, or the caller's stream) and decodes vertex, edge and neighbour;
//    a window never crosses a batch, so the lane holding a step in the batch evaluates it;
//  * each lane reads the flavours and labels at its two sites from the LDS site map;
//  * RES: every wavefront takes the steps whose moving row it owns (a ballot mask), picks
//    q_up or q_dn out of its register tables (indexed moves) and publishes it in LDS; one
//    barrier later each lane holds its step's ratio (RES = false: the lane loads them);
//  * every lane applies MetroMC::YesNo to its step; a ballot gives the first accepted
//    move.  The steps before it are finished (rejected or trivial), the accepted step
//    applies the rank-1 update (its rows go through LDS, one more barrier), and the next
//    window starts after it.  Without an accepted move the window simply advances.
// Every decision is the one MetroMC::Step takes on the same uniforms, in the same order.
// slots per pipelined chunk of a streamed row: the largest divisor of the lane's slot count
// that is at most 12
constexpr int stream_chunk(int cpt)
{
    for (int d = 12; d > 1; --d)
        if (cpt % d == 0) return d;
    return 1;
}

template <typename T, int N, int LPR, bool RES, bool MONLY, bool WDU = false, bool TR = true, bool MF = false>
__global__ void __launch_bounds__(((N * LPR + 63) / 64) * 64)
sweep_kernel(SweepArgs a)
{
    // TR = false: no replay stream and no traces (the production launch); the kernel then
    // carries none of their pointers or tests
    const double *const uniforms = TR ? a.uniforms : nullptr;
    int8_t *const tr_noop = TR ? a.tr_noop : nullptr;
    int8_t *const tr_accept = TR ? a.tr_accept : nullptr;
    double *const tr_ratio = TR ? a.tr_ratio : nullptr;
    double *const tr_eloc = TR ? a.tr_eloc : nullptr;
    double *const tr_mz = TR ? a.tr_mz : nullptr;
    double *const tr_logw = TR ? a.tr_logw : nullptr;
    // steps of this launch (the host splits longer walks) and the measurement interval
    const int n_steps = (int)a.n_steps;
    const int meas = (int)a.meas;
    using C = typename cpx<T>::t;
    constexpr int NT = ((N * LPR + 63) / 64) * 64;
    constexpr int CPT = N / LPR;
    constexpr int NS = 2 * N;
    constexpr int RPW = 64 / LPR;                    // table rows per wavefront
    constexpr int KW = 16;                           // steps evaluated per window
    constexpr int RP = CPT + 16 / (int)sizeof(T);    // LDS row pitch: one 16-byte pad
    // slot loops: fully unrolled over register tables (static register indices), partly
    // over streamed ones (so the loads of a long row are not all hoisted at once)
    constexpr int UNR = RES ? CPT : 4;
    static_assert(N % LPR == 0, "LPR must divide N");
    static_assert(64 % LPR == 0, "row groups must not straddle wavefronts");
    static_assert(NT <= 1024, "workgroup too large");
    // Bond ratio cache (register tables, one-move windows): for every lattice bond b (edge
    // b of a.edge) whose two sites hold an up particle i and a down particle j, s_bq[b]
    // holds (G_up[i][j], G_dn[j][i]), the two single-hop ratios of swapping them.  It is
    // updated with the table's own rank-1 arithmetic after every accepted move, so each
    // entry equals the table entry bit for bit; proposals and ProjHeis read it instead of
    // the register tables.
    constexpr bool BQ = RES && !WDU && !MONLY;
    // MF (FP32, N = 128, 512 threads): the tables live in the accumulator layout of
    // v_mfma_f32_32x32x2_f32 and each accepted move's rank-1 update is one K = 2 matrix
    // operation per 32 x 32 tile.  A complex rank-1 update is a real rank-2 update of each
    // real part: Re G += [Re nc, -Im nc] [Re v; Im v], Im G += [Im nc, Re nc] [Re v; Im v],
    // with nc the row's coefficient and v the moving row.  Wavefront w holds row block
    // w / 2 and column blocks 2 (w % 2) and 2 (w % 2) + 1 of both parts of both tables; lane
    // l of a tile holds column l % 32 and, in register r, row 8 (r / 4) + 4 (l / 32) + r % 4.
    // The bond cache follows the tables to FP32 rounding (its update is the VALU sequence).
    static_assert(!MF || (N == 128 && LPR == 4 && sizeof(T) == 4 && BQ), "matrix-core tables: FP32, N = 128");
    constexpr int NB = 2 * NS;                       // bonds of the periodic square lattice
    constexpr int NBT = BQ ? (NB + NT - 1) / NT : 1; // bonds per thread
    static_assert(!BQ || NS <= 1024, "bond cache packs v1, v2 in 10 bits each");

    __shared__ int s_site[NS];        // (flavour << 16) | label
    __shared__ int s_part[2][N];      // label -> site
    __shared__ int s_nbr[NS * 4];     // neighbour site; BQ: | (bond index << 16)
    __shared__ C s_q[BQ ? 1 : 2][64][2];  // [parity][lane of the step][flavour] single-hop ratios
    // [flavour][lane of the row group][re, im][slot]: the moving row, e_cs added
    __shared__ __align__(16) T s_row[2][LPR][2][RP];
    __shared__ C s_qb[BQ ? 1 : NS * 4];   // measurement: G entry of the particle at site s for neighbour d
    __shared__ __align__(16) C s_bq[BQ ? NB : 1][2];
    // BQ, accepted move: per flavour and row, the update coefficient -(G[row][cs] - delta)/q
    // and the column entry G[row][cs] - delta (cs = cj for the up table, ci for the down one)
    __shared__ C s_nc[BQ ? 2 : 1][BQ ? N : 1];
    __shared__ C s_ux[BQ ? 2 : 1][BQ ? N : 1];
    __shared__ double s_red[NT / 64][4];
    __shared__ double s_acc[ACC_STRIDE];
    // amplitude (s_amp) at each measurement since the last flush, its trace slot and
    // whether it is accumulated; wavefront 0 forms their log-weights AR at a time
    constexpr int AR = 64;
    __shared__ double s_snap[AR][6];
    __shared__ int s_snap_slot[AR];
    constexpr int LS = isqrt_c(NS);  // lattice side
    static_assert(LS * LS == NS, "square lattice");
    // thread 0: amplitude change per flavour as complex mantissa [0..3] and binary
    // exponent [4..5], so no logarithm is taken per step
    __shared__ double s_amp[6];
    // StatSpinStruct: a chunk of the transposed G_dn (RES kernels)
    constexpr int TRK = LPR * N * 2 * (int)sizeof(T);         // bytes per slot k of G_dn^T
    constexpr int CH = divisor_within(CPT, TRK, 32768);         // slots per chunk
    __shared__ double s_tr[(RES && !WDU) ? (CH * TRK + 7) / 8 : 1];
    // WDU: row dumps and window scratch (dynamic), which the StatSpinStruct chunks reuse
    extern __shared__ __align__(16) char smem[];
    static_assert(!WDU || (RES && N <= 128 && !MONLY), "delayed-update windows need register tables, N <= 128");
    static_assert(!WDU || (size_t)CH * TRK <= 2 * wdu_dump_elems<T, N, LPR>() * sizeof(T), "statspin chunk");

    const int w = __builtin_amdgcn_readfirstlane(a.order ? a.order[blockIdx.x] : (int)blockIdx.x);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int row = tid / LPR, c = tid % LPR;
    const int gbase = lane & ~(LPR - 1);
    const bool has_row = row < N;

    for (int i = tid; i < NS; i += NT) s_site[i] = a.site[(size_t)w * NS + i];
    for (int i = tid; i < 2 * N; i += NT) s_part[i / N][i % N] = a.part[(size_t)w * 2 * N + i];
    for (int i = tid; i < NS * 4; i += NT) s_nbr[i] = a.nbr[i];
    if (tid < ACC_STRIDE) s_acc[tid] = 0.0;
    if (tid == 0) {
        s_amp[0] = 1.0; s_amp[1] = 0.0; s_amp[2] = 1.0; s_amp[3] = 0.0;
        s_amp[4] = 0.0; s_amp[5] = 0.0;
    }
    int n_noop = 0, n_acc = 0, n_rej = 0;

    C *gG = reinterpret_cast<C *>(a.G) + (size_t)w * 2 * N * N;
    auto gat = [&](int fl, int i, int j) -> C & { return gG[(size_t)fl * N * N + (size_t)i * N + j]; };

    // RES: real and imaginary parts of the thread's CPT slots of each table (the
    // measurement-only pass reads just the bond entries it needs from HBM instead)
    regvec<T, CPT> gr[2], gi[2];
    // MF: tiles [flavour][column block of the pair] of the real and imaginary parts
    regvec<float, 16> mr[2][2], mi[2][2];
    const int mrb = wid >> 1, mcp = wid & 1, mh = lane >> 5, ml = lane & 31;
    if constexpr (MF) {
#pragma unroll
        for (int fl = 0; fl < 2; ++fl)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const C e = gat(fl, mrb * 32 + (r >> 2) * 8 + mh * 4 + (r & 3), (2 * mcp + cb) * 32 + ml);
                    mr[fl][cb][r] = e.x;
                    mi[fl][cb][r] = e.y;
                }
    } else if constexpr (RES && !MONLY) {
        if (has_row) {
#pragma unroll
            for (int fl = 0; fl < 2; ++fl)
#pragma unroll UNR
                for (int k = 0; k < CPT; ++k) {
                    const C e = gat(fl, row, c + LPR * k);
                    gr[fl][k] = e.x;
                    gi[fl][k] = e.y;
                }
        }
    }
    const uint64_t ctr0 = a.counter[w];
    // ln|det| of both flavours at the launch's start (the launch adds its ratios at the end)
    const double la0 = a.logabs[2 * w] + a.logabs[2 * w + 1];
    int mslot = 0;
    __syncthreads();

    // BQ: bond b = tid + m NT of this thread, its end sites (a | t << 16) and their site
    // words, kept current move by move
    int bend[NBT], binfo[NBT][2];
    if constexpr (BQ) {
#pragma unroll
        for (int m = 0; m < NBT; ++m) {
            const int b = tid + m * NT;
            bend[m] = 0;
            binfo[m][0] = binfo[m][1] = 0;
            if (b < NB) {
                const int4 eg = a.edge[b];
                s_nbr[eg.x * 4 + eg.y] |= b << 16;
                s_nbr[eg.z * 4 + eg.w] |= b << 16;
                bend[m] = eg.x | (eg.z << 16);
                binfo[m][0] = s_site[eg.x];
                binfo[m][1] = s_site[eg.z];
            }
        }
        __syncthreads();
        // the owner of row i of G_fl writes G_fl[i][j] for every neighbour site holding a
        // particle j of the other flavour (MF: every bond thread reads its two entries from
        // the tables in global memory)
        if constexpr (MF) {
#pragma unroll
            for (int m = 0; m < NBT; ++m) {
                const int ia = binfo[m][0], it = binfo[m][1];
                if (tid + m * NT < NB && ((ia ^ it) >> 16)) {
                    const bool aup = (ia >> 16) == 0;
                    const int i = (aup ? ia : it) & 0xFFFF, jd = (aup ? it : ia) & 0xFFFF;
                    s_bq[tid + m * NT][0] = gat(0, i, jd);
                    s_bq[tid + m * NT][1] = gat(1, jd, i);
                }
            }
        } else if (has_row) {
#pragma unroll
            for (int fl = 0; fl < 2; ++fl) {
                const int sh = s_part[fl][row];
                for (int d = 0; d < 4; ++d) {
                    const int nb = s_nbr[sh * 4 + d];
                    const int info = s_site[nb & 0xFFFF];
                    if ((info >> 16) == fl) continue;  // parallel bond
                    const int j = info & 0xFFFF;
                    if (c == j % LPR) {
                        const int kj = j / LPR;
                        C val = cmk<C>(0, 0);
#pragma unroll UNR
                        for (int k = 0; k < CPT; ++k) {
                            const T m = (k == kj) ? (T)1 : (T)0;
                            val.x = fma(m, gr[fl][k], val.x);
                            val.y = fma(m, gi[fl][k], val.y);
                        }
                        s_bq[nb >> 16][fl] = val;
                    }
                }
            }
        }
        __syncthreads();
    }

    // StatSpinStruct::measure (StatSpinStruct.cpp:16-135) into this walker's FP64 sums:
    // szz[x][y] += sz_x sz_y (1/4 each); spm or smp diagonal += 1 for up or down at x; for
    // every up particle i (site a) and down particle j (site b) smp[a][b] -= R_ij with the
    // swap ratio R_ij = G_up[i][j] G_dn[j][i] (conj(amp) amp_swap / w of the reference).
    // G_dn goes through LDS transposed, in chunks of CH slots of every row group.
    auto statspin = [&]() {
        double *ss = a.ss + (size_t)w * ss_block(NS);
        double *zz = ss;
        double *mp = ss + (size_t)NS * NS;
        double *pd = ss + (size_t)3 * NS * NS;
        for (int e = tid; e < NS * NS; e += NT) {
            const int x = e / NS, y = e - x * NS;
            zz[e] += ((s_site[x] >> 16) == (s_site[y] >> 16)) ? 0.25 : -0.25;
        }
        for (int v = tid; v < NS; v += NT) {
            if ((s_site[v] >> 16) == 0) pd[v] += 1.0;
            else mp[2 * ((size_t)v * NS + v)] += 1.0;
        }
        if (tid == 0) pd[NS] += 1.0;
        const int sa = has_row ? s_part[0][row] : 0;
#pragma unroll
        for (int k0 = 0; k0 < CPT; k0 += CH) {
            // rows j = c' + LPR k of G_dn, k in [k0, k0 + CH), i.e. j in [LPR k0, LPR (k0 + CH))
            if constexpr (RES) {
                T *sT = WDU ? reinterpret_cast<T *>(smem) : reinterpret_cast<T *>(s_tr);
                __syncthreads();
                if (has_row && row >= LPR * k0 && row < LPR * (k0 + CH)) {
#pragma unroll UNR
                    for (int k = 0; k < CPT; ++k) {
                        const int i = c + LPR * k;
                        const size_t o = ((size_t)(row - LPR * k0) * N + i) * 2;
                        sT[o] = gr[1][k];
                        sT[o + 1] = gi[1][k];
                    }
                }
                __syncthreads();
                if (has_row) {
#pragma unroll UNR
                    for (int k = k0; k < k0 + CH; ++k) {
                        const int j = c + LPR * k;
                        const size_t o = ((size_t)(j - LPR * k0) * N + row) * 2;
                        const double br = sT[o], bi = sT[o + 1];
                        const double ar = gr[0][k], ai = gi[0][k];
                        double *m = mp + 2 * ((size_t)sa * NS + s_part[1][j]);
                        m[0] -= ar * br - ai * bi;
                        m[1] -= ar * bi + ai * br;
                    }
                }
            } else {
                if (has_row) {
#pragma unroll UNR
                    for (int k = k0; k < k0 + CH; ++k) {
                        const int j = c + LPR * k;
                        const C ga = gat(0, row, j), gb = gat(1, j, row);
                        double *m = mp + 2 * ((size_t)sa * NS + s_part[1][j]);
                        m[0] -= (double)ga.x * gb.x - (double)ga.y * gb.y;
                        m[1] -= (double)ga.x * gb.y + (double)ga.y * gb.x;
                    }
                }
            }
        }
    };

    // log-weights of the snapshots, ln |amp|^2 = 2 (ln|det_up| + ln|det_dn|), one per lane of
    // wavefront 0, into the log-weight trace and (in measurement order) s_acc[11]
    int nsnap = 0;
    auto flush_snaps = [&]() {
        if (wid == 0 && nsnap > 0) {
            double lw = 0.0;
            if (lane < nsnap) {
                const double *sp = s_snap[lane];
                const double la = la0 + log(hypot(sp[0], sp[1])) + log(hypot(sp[2], sp[3])) +
                                  (sp[4] + sp[5]) * 0.69314718055994531;
                lw = 2.0 * la;
                const int slot = s_snap_slot[lane] & 0x7FFFFFFF;
                if (tr_logw) tr_logw[(size_t)w * a.n_meas_slots + slot] = lw;
            }
            double t = s_acc[11];
            for (int i = 0; i < nsnap; ++i) {
                const double v = readlane_real(lw, i);
                if (s_snap_slot[i] >= 0) t += v;
            }
            if (lane == 0) s_acc[11] = t;
        }
        nsnap = 0;
    };

    // Quantity::measure (MetroMC.cpp:61-68); s < 0 is the measurement-only pass
    auto measure = [&](int s) {
        __syncthreads();
        double ev[2] = {0.0, 0.0};
        if (BQ && (a.mask & MEAS_PROJHEIS)) {
            // ProjHeis.cpp:131-166 over the bond cache; qa is the ratio of the particle at
            // the edge's first site
#pragma unroll
            for (int m = 0; m < NBT; ++m) {
                if (tid + m * NT < NB) {
                    const int ia = binfo[m][0], it = binfo[m][1];
                    if (((ia ^ it) >> 16) == 0) {
                        ev[0] += 0.25;
                    } else {
                        const int fa = ia >> 16;
                        const C qa = s_bq[tid + m * NT][fa], qb = s_bq[tid + m * NT][fa ^ 1];
                        const double pr = (double)qa.x * qb.x - (double)qa.y * qb.y;
                        const double pi = (double)qa.x * qb.y + (double)qa.y * qb.x;
                        ev[0] += -0.25 - 0.5 * pr;
                        ev[1] += -0.5 * pi;
                    }
                }
            }
        } else if (a.mask & MEAS_PROJHEIS) {
            if (has_row) {
#pragma unroll
                for (int fl = 0; fl < 2; ++fl) {
                    const int sh = s_part[fl][row];
                    for (int d = 0; d < 4; ++d) {
                        const int t = s_nbr[sh * 4 + d] & 0xFFFF;
                        const int info = s_site[t];
                        if ((info >> 16) == fl) continue;  // parallel bond
                        const int j = info & 0xFFFF;
                        if (c == j % LPR) {
                            if constexpr (RES && !MONLY) {
                                // masked sum over the slots (the slot index differs
                                // between lanes): no run-time register index
                                const int kj = j / LPR;
                                C val = cmk<C>(0, 0);
#pragma unroll UNR
                                for (int k = 0; k < CPT; ++k) {
                                    const T m = (k == kj) ? (T)1 : (T)0;
                                    val.x = fma(m, gr[fl][k], val.x);
                                    val.y = fma(m, gi[fl][k], val.y);
                                }
                                s_qb[sh * 4 + d] = val;
                            } else {
                                s_qb[sh * 4 + d] = gat(fl, row, j);
                            }
                        }
                    }
                }
            }
            __syncthreads();
            // ProjHeis.cpp:131-166: -(1/4 + 1/2 q_up q_dn) per antiparallel edge, +1/4 per parallel
            for (int e = tid; e < 2 * NS; e += NT) {
                const int4 eg = a.edge[e];
                if ((s_site[eg.x] >> 16) == (s_site[eg.z] >> 16)) {
                    ev[0] += 0.25;
                } else {
                    const C qa = s_qb[eg.x * 4 + eg.y], qb = s_qb[eg.z * 4 + eg.w];
                    const double pr = (double)qa.x * qb.x - (double)qa.y * qb.y;
                    const double pi = (double)qa.x * qb.y + (double)qa.y * qb.x;
                    ev[0] += -0.25 - 0.5 * pr;
                    ev[1] += -0.5 * pi;
                }
            }
        }
        if ((a.mask & MEAS_STATSPIN) && s >= 0) statspin();
        double mv2 = 0.0;
        if (a.mask & MEAS_STAGMAGNZ) {
            // StagMagnZ.cpp:12-43: (1/Ns) sum_v (-1)^(x+y) (2 up_v - 1)
            for (int v = tid; v < NS; v += NT) {
                const int x = v % LS, y = v / LS;
                const double sg = ((x + y) & 1) ? -1.0 : 1.0;
                mv2 += sg * (((s_site[v] >> 16) == 0) ? 1.0 : -1.0);
            }
        }
        // the amplitude now (read before the next barrier: an accepted move after this
        // measurement rewrites s_amp only behind it)
        if (tid == 0) {
#pragma unroll
            for (int i = 0; i < 6; ++i) s_snap[nsnap][i] = s_amp[i];
            s_snap_slot[nsnap] = mslot | (s >= 0 ? 0 : (int)0x80000000);
        }
        // per-wavefront sums (DPP, lane 63), then one barrier; thread 0 adds the partials
        double red[3] = {ev[0], ev[1], mv2};
#pragma unroll
        for (int i = 0; i < 3; ++i) red[i] = wave_sum_to_lane63(red[i]);
        if (lane == 63) {
#pragma unroll
            for (int i = 0; i < 3; ++i) s_red[wid][i] = red[i];
        }
        __syncthreads();
        if (tid == 0) {
            double t[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < NT / 64; ++q)
#pragma unroll
                for (int i = 0; i < 3; ++i) t[i] += s_red[q][i];
            const double E_re = t[0] / NS, E_im = t[1] / NS, mz = t[2] / NS;
            if (s >= 0) {
                // explicit FMAs: the same rounding in every instantiation of the kernel
                if (a.mask & MEAS_PROJHEIS) {
                    s_acc[0] += E_re;
                    s_acc[1] += E_im;
                    s_acc[2] += fma(E_re, E_re, E_im * E_im);
                }
                if (a.mask & MEAS_STAGMAGNZ) {
                    s_acc[4] += mz;
                    s_acc[6] = fma(mz, mz, s_acc[6]);
                }
                s_acc[3] += 1;
            }
            const size_t mi = (size_t)w * a.n_meas_slots + mslot;
            if (tr_eloc) { tr_eloc[2 * mi] = E_re; tr_eloc[2 * mi + 1] = E_im; }
            if (tr_mz) { tr_mz[2 * mi] = mz; tr_mz[2 * mi + 1] = 0.0; }
        }
        ++mslot;
        if (++nsnap == AR) flush_snaps();
        // no closing barrier: thread 0 alone touches s_acc, s_red and the snapshots until
        // the next measurement's barriers
    };

    // uniforms of the batch of 64 steps starting at s0 (s0 % 64 == 0): lane l decodes step
    // s0 + l into bpv = v1 | (v2 << 16) and its acceptance uniform bu3
    auto draw_batch = [&](int s0, int &bpv, double &bu3) {
        const int sl = s0 + lane;
        int v1 = 0, ed = 0;
        double u3 = 1.0;
        if (uniforms) {
            if (sl < n_steps) {
                const double *up = uniforms + ((size_t)w * n_steps + sl) * 4;
                v1 = (int)(NS * up[0]);
                v1 = v1 < NS ? v1 : NS - 1;
                ed = (int)(4 * up[1]);
                ed = ed < 4 ? ed : 3;
                u3 = up[3];
            }
        } else {
            const uint64_t st = ctr0 + (uint64_t)sl;
            uint32_t cc[4] = {(uint32_t)st, (uint32_t)(st >> 32), (uint32_t)w, 0u};
            // the key is made opaque here so that its ten round keys are formed per draw
            // rather than held in scalar registers through the whole walk
            uint32_t k0 = a.key0, k1 = a.key1;
            asm volatile("" : "+s"(k0), "+s"(k1));
            philox4x32_10_dev(cc, k0, k1);
            // floor(NS u0) and floor(4 u1) of u = (x + 1/2) 2^-32, in integers
            // (exact: NS (2x + 1) < 2^42); cc[2] is the internal-state draw, unused
            v1 = (int)(((uint64_t)NS * (2ull * cc[0] + 1ull)) >> 33);
            ed = (int)(cc[1] >> 30);
            u3 = u32_to_open01(cc[3]);
        }
        const int nb = s_nbr[v1 * 4 + ed];
        // BQ: v1 | v2 << 10 | bond << 20; otherwise v1 | v2 << 16
        bpv = BQ ? (v1 | ((nb & 0xFFFF) << 10) | ((nb >> 16) << 20)) : (v1 | (nb << 16));
        bu3 = u3;
    };

    if constexpr (MONLY) {
        measure(-1);
    } else if constexpr (!WDU) {
        int bpv = 0;          // batch: v1 | (v2 << 16) of step (s0 & ~63) + lane
        double bu3 = 1.0;     // batch: acceptance uniform of that step
        int par = 0;
        int s0 = 0;
        // the next measured step, meas > 0: measurements follow step s when meas == 1 or
        // s % meas == 0, and no window runs past one
        int next_meas = 0;
        while (s0 < n_steps) {
            const int b0 = (int)(s0 & 63);
            if (b0 == 0) draw_batch(s0, bpv, bu3);
            // window: up to KW steps, within the batch, ending at the next measured step
            int lim = n_steps - s0;
            if (lim > 64 - b0) lim = 64 - b0;
            if (meas > 0 && next_meas - s0 + 1 < lim) lim = next_meas - s0 + 1;
            const int len = (int)(lim < KW ? lim : KW);
            // lane b0 + i holds step s0 + i of the window, as it holds it in the batch
            const bool inw = lane >= b0 && lane < b0 + len;
            const int v1 = BQ ? (bpv & 0x3FF) : (bpv & 0xFFFF);
            const int v2 = BQ ? ((bpv >> 10) & 0x3FF) : (bpv >> 16);
            const double u3 = bu3;
            // the window's lanes read the two site words (BQ: and the bond's cached ratios);
            // the other lanes see a trivial step, which no count or decision takes
            int i1 = 0, i2 = 0;
            C q0l = cmk<C>(0, 0), q1l = cmk<C>(0, 0);
            if (inw) {
                i1 = s_site[v1];
                i2 = s_site[v2];
                if constexpr (BQ) {
                    q0l = s_bq[bpv >> 20][0];
                    q1l = s_bq[bpv >> 20][1];
                }
            }
            const bool noop = ((i1 ^ i2) >> 16) == 0;
            // up particle ci on site sa, down particle cj on site sb
            const bool up1 = (i1 >> 16) == 0;
            const int ci = (up1 ? i1 : i2) & 0xFFFF, cj = (up1 ? i2 : i1) & 0xFFFF;
            const int sa = up1 ? v1 : v2, sb = up1 ? v2 : v1;
            if constexpr (BQ) {
                // the bond's cached ratios were read with the site words (used only on
                // antiparallel bonds); the cache is current for this window: the last accepted
                // move ended with a barrier after its update
            } else if constexpr (RES) {
                // each wavefront publishes the ratios whose rows it owns: q_up of step i
                // sits in row ci of G_up, q_dn in row cj of G_dn
                uint64_t own = __ballot(inw && !noop && ci / RPW == wid);
                while (own) {
                    const int i = (int)__builtin_ctzll(own);
                    own &= own - 1;
                    const int ci_i = __builtin_amdgcn_readlane(ci, i);
                    const int cj_i = __builtin_amdgcn_readlane(cj, i);
                    const int k = cj_i / LPR;
                    const C x = cmk<C>(gr[0][k], gi[0][k]);
                    if (row == ci_i && c == cj_i % LPR) s_q[par][i][0] = x;
                }
                own = __ballot(inw && !noop && cj / RPW == wid);
                while (own) {
                    const int i = (int)__builtin_ctzll(own);
                    own &= own - 1;
                    const int ci_i = __builtin_amdgcn_readlane(ci, i);
                    const int cj_i = __builtin_amdgcn_readlane(cj, i);
                    const int k = ci_i / LPR;
                    const C x = cmk<C>(gr[1][k], gi[1][k]);
                    if (row == cj_i && c == ci_i % LPR) s_q[par][i][1] = x;
                }
                __syncthreads();
                if (inw && !noop) {
                    q0l = s_q[par][lane][0];
                    q1l = s_q[par][lane][1];
                }
            } else {
                if (inw && !noop) {
                    q0l = gat(0, ci, cj);
                    q1l = gat(1, cj, ci);
                }
            }
            double ratio = 1.0;
            bool accl;
            if (noop) {
                accl = u3 < 1.0;  // LatticeStepper.cpp:64-68 returns the current weight: ratio 1
            } else {
                const double n0 = (double)q0l.x * q0l.x + (double)q0l.y * q0l.y;
                const double n1 = (double)q1l.x * q1l.x + (double)q1l.y * q1l.y;
                ratio = n0 * n1;
                accl = ratio > 1.0 ? (u3 < 1.0) : (u3 < ratio);  // MetroMC::YesNo, MetroMC.cpp:27-43
            }
            // first accepted move of the window: lane j, step s0 + j - b0
            const uint64_t moved = __ballot(inw && !noop && accl);
            const int j = moved ? (int)__builtin_ctzll(moved) : 64;
            const int nproc = moved ? j - b0 + 1 : len;
            const uint64_t pmask = (nproc >= 64 ? ~0ull : ((1ull << nproc) - 1ull)) << b0;
            const int nnoop = __popcll(__ballot(noop) & pmask);
            const int nmove = moved ? 1 : 0;
            n_noop += nnoop;
            n_acc += nmove;
            n_rej += nproc - nnoop - nmove;
            if (TR && wid == 0 && ((pmask >> lane) & 1ull)) {
                const size_t ti = (size_t)w * n_steps + s0 + (lane - b0);
                if (tr_noop) tr_noop[ti] = noop ? 1 : 0;
                if (tr_accept) tr_accept[ti] = accl ? 1 : 0;
                if (tr_ratio) tr_ratio[ti] = ratio;
            }
            // The accepted move is wave-uniform; taken through an opaque copy it is
            // compiled as an exec-masked region (skipped by s_cbranch_execz like a scalar
            // branch), which keeps the tables in the same registers on both paths.
            int jv;
            asm volatile("v_mov_b32 %0, %1" : "=v"(jv) : "s"(j));
            if (jv < 64) {
                const int ci_j = __builtin_amdgcn_readlane(ci, j), cj_j = __builtin_amdgcn_readlane(cj, j);
                const int sa_j = __builtin_amdgcn_readlane(sa, j), sb_j = __builtin_amdgcn_readlane(sb, j);
                const C q0 = cmk<C>(readlane_real(q0l.x, j), readlane_real(q0l.y, j));
                const C q1 = cmk<C>(readlane_real(q1l.x, j), readlane_real(q1l.y, j));
                const double n0 = (double)q0.x * q0.x + (double)q0.y * q0.y;
                const double n1 = (double)q1.x * q1.x + (double)q1.y * q1.y;
                // 1/q per flavour, one reciprocal each, formed while the moving rows are
                // stashed (kept ahead of the barrier below)
                double rn0, rn1;
                if constexpr (sizeof(T) == 4) {
                    // the coefficients are rounded to FP32: FP64 reciprocal estimate and two
                    // Newton steps (full FP64 accuracy before that rounding)
                    rn0 = __builtin_amdgcn_rcp(n0);
                    rn1 = __builtin_amdgcn_rcp(n1);
                    rn0 = fma(rn0, fma(-n0, rn0, 1.0), rn0);
                    rn1 = fma(rn1, fma(-n1, rn1, 1.0), rn1);
                    rn0 = fma(rn0, fma(-n0, rn0, 1.0), rn0);
                    rn1 = fma(rn1, fma(-n1, rn1, 1.0), rn1);
                } else {
                    rn0 = 1.0 / n0;
                    rn1 = 1.0 / n1;
                }
                const C qi0 = cmk<C>(q0.x * rn0, -q0.y * rn0), qi1 = cmk<C>(q1.x * rn1, -q1.y * rn1);
                // column cj of G_up and column ci of G_dn: slot ks, lane ls of each row group
                const int ks0 = cj_j / LPR, ls0 = cj_j % LPR, ks1 = ci_j / LPR, ls1 = ci_j % LPR;
                C x0 = cmk<C>(0, 0), x1 = cmk<C>(0, 0);
                if constexpr (MF) {
                } else if constexpr (RES) {
                    x0 = cmk<C>(gr[0][ks0], gi[0][ks0]);
                    x1 = cmk<C>(gr[1][ks1], gi[1][ks1]);
                } else {
                    x0 = has_row ? gat(0, row, cj_j) : cmk<C>(0, 0);
                    x1 = has_row ? gat(1, row, ci_j) : cmk<C>(0, 0);
                }
                // u_i = G[i][cs] - delta(i, rs), held by lane ls of row i's group, and the
                // row's update coefficient -u_i / q
                auto coefficient = [&](int fl, C &uu) -> C {
                    const C qi = fl ? qi1 : qi0;
                    const C x = fl ? x1 : x0;
                    const int ls = fl ? ls1 : ls0, rsel = fl ? cj_j : ci_j;
                    uu.x = __shfl(x.x, gbase + ls);
                    uu.y = __shfl(x.y, gbase + ls);
                    if (row == rsel) uu.x -= (T)1;
                    const C coef = cmul(uu, qi);
                    return cmk<C>(-coef.x, -coef.y);
                };
                // BQ: both coefficients before the barrier, published with u for the bond
                // cache update
                C ncb[2];
                // MF: the moving columns G_up[:, cj] and G_dn[:, ci] (raw entries, re | im by row)
                // and the moving rows G_up[ci, :] + e_cj and G_dn[cj, :] + e_ci (re | im by column)
                T *const mux = reinterpret_cast<T *>(&s_ux[0][0]);
                T *const mrow = &s_row[0][0][0][0];
                // MF: the row coefficient -(G[row][cs] - delta) / q from the published column
                // (formed after the barrier by the table update and the bond cache alike)
                auto mf_coefficient = [&](int fl, int i) -> C {
                    C uu = cmk<C>(mux[(fl * 2) * N + i], mux[(fl * 2 + 1) * N + i]);
                    if (i == (fl ? cj_j : ci_j)) uu.x -= (T)1;
                    const C coef = cmul(uu, fl ? qi1 : qi0);
                    return cmk<C>(-coef.x, -coef.y);
                };
                if constexpr (MF) {
#pragma unroll
                    for (int fl = 0; fl < 2; ++fl) {
                        const int cs = fl ? ci_j : cj_j, rs = fl ? cj_j : ci_j;
                        // column: the lanes of column cs % 32 in the wavefronts holding its block
                        const int cbc = cs >> 5;
                        if ((cbc >> 1) == mcp && ml == (cs & 31)) {
#pragma unroll
                            for (int cb = 0; cb < 2; ++cb) {
                                if ((cbc & 1) == cb) {
#pragma unroll
                                    for (int q = 0; q < 4; ++q) {
                                        const int r0 = mrb * 32 + q * 8 + mh * 4;
                                        *reinterpret_cast<float4 *>(mux + (fl * 2) * N + r0) =
                                            make_float4(mr[fl][cb][4 * q], mr[fl][cb][4 * q + 1],
                                                        mr[fl][cb][4 * q + 2], mr[fl][cb][4 * q + 3]);
                                        *reinterpret_cast<float4 *>(mux + (fl * 2 + 1) * N + r0) =
                                            make_float4(mi[fl][cb][4 * q], mi[fl][cb][4 * q + 1],
                                                        mi[fl][cb][4 * q + 2], mi[fl][cb][4 * q + 3]);
                                    }
                                }
                            }
                        }
                        // row: register 4 (rr / 8) + rr % 4 of the lanes in half (rr / 4) % 2
                        const int rr = rs & 31;
                        if (mrb == (rs >> 5) && mh == ((rr >> 2) & 1)) {
                            const int reg = (rr >> 3) * 4 + (rr & 3);
#pragma unroll
                            for (int cb = 0; cb < 2; ++cb) {
                                const int col = (2 * mcp + cb) * 32 + ml;
                                T vr = mr[fl][cb][reg];
                                const T vi = mi[fl][cb][reg];
                                if (col == cs) vr += (T)1;
                                mrow[(fl * 2) * N + col] = vr;
                                mrow[(fl * 2 + 1) * N + col] = vi;
                            }
                        }
                    }
                } else if constexpr (BQ) {
#pragma unroll
                    for (int fl = 0; fl < 2; ++fl) {
                        C uu;
                        ncb[fl] = coefficient(fl, uu);
                        if (has_row && c == 0) {
                            s_nc[fl][row] = ncb[fl];
                            s_ux[fl][row] = uu;
                        }
                    }
                }
                // the moving rows G_up[ci,:] + e_cj and G_dn[cj,:] + e_ci go to LDS
#pragma unroll
                for (int fl = 0; fl < 2; ++fl) {
                    const int rsel = fl ? cj_j : ci_j;
                    if (!MF && row == rsel) {
                        T *dr = &s_row[fl][c][0][0], *di = &s_row[fl][c][1][0];
                        if constexpr (RES) {
#pragma unroll
                            for (int k = 0; k < CPT; ++k) {
                                dr[k] = gr[fl][k];
                                di[k] = gi[fl][k];
                            }
                        } else {
                            // streamed row in chunks, the next chunk's loads in flight
                            constexpr int SK = stream_chunk(CPT), NCH = CPT / SK;
                            C ea[SK], eb[SK];
                            auto ldk = [&](int k0, C (&e)[SK]) {
#pragma unroll
                                for (int j = 0; j < SK; ++j) e[j] = gat(fl, row, c + LPR * (k0 + j));
                            };
                            auto stk = [&](int k0, const C (&e)[SK]) {
#pragma unroll
                                for (int j = 0; j < SK; ++j) {
                                    dr[k0 + j] = e[j].x;
                                    di[k0 + j] = e[j].y;
                                }
                            };
                            ldk(0, ea);
#pragma unroll
                            for (int ch = 0; ch < NCH; ch += 2) {
                                if (ch + 1 < NCH) ldk((ch + 1) * SK, eb);
                                stk(ch * SK, ea);
                                if (ch + 2 < NCH) ldk((ch + 2) * SK, ea);
                                if (ch + 1 < NCH) stk((ch + 1) * SK, eb);
                            }
                        }
                        if (c == (fl ? ls1 : ls0)) {
                            const C x = fl ? x1 : x0;
                            dr[fl ? ks1 : ks0] = x.x + (T)1;
                        }
                    }
                }
                // State::Hop (State.cpp:73-84) and amplitude *= q per flavour
                // (SlaterDeterminant.cpp:321)
                auto hop = [&]() {
                    s_site[sb_j] = ci_j;
                    s_site[sa_j] = (1 << 16) | cj_j;
                    s_part[0][ci_j] = sb_j;
                    s_part[1][cj_j] = sa_j;
                };
                auto amplitude = [&]() {
#pragma unroll
                    for (int fl = 0; fl < 2; ++fl) {
                        const C q = fl ? q1 : q0;
                        double mr = s_amp[2 * fl], mi = s_amp[2 * fl + 1], ex = s_amp[4 + fl];
                        amp_mul(mr, mi, ex, (double)q.x, (double)q.y);
                        s_amp[2 * fl] = mr;
                        s_amp[2 * fl + 1] = mi;
                        s_amp[4 + fl] = ex;
                    }
                };
                // BQ: one wavefront that stashes no row updates the amplitude before the
                // barrier and the site map after it (every wavefront reads the window's sites
                // for itself, so the map changes only once all of them are past the barrier;
                // the next reads follow the barrier that ends the move)
                constexpr int NW = NT / 64;
                const int wu = ci_j / RPW, wd = cj_j / RPW;
                const int wb = NW < 3 ? 0 : ((wu != 0 && wd != 0) ? 0 : ((wu != 1 && wd != 1) ? 1 : 2));
                if constexpr (BQ) {
                    if (wid == wb && lane == 0) amplitude();
                }
                asm volatile("" ::"v"(qi0.x), "v"(qi0.y), "v"(qi1.x), "v"(qi1.y));
                __syncthreads();
                if constexpr (BQ) {
                    if (wid == wb && lane == 0) hop();
                }
                auto bond_update = [&]() {
                    // bond cache: G'[i][j] = G[i][j] + nc_i (G[rs][j] + delta(j, cs)) in the
                    // table update's operation order.  A bond that holds an antiparallel pair
                    // only after this move holds ci or cj: its old entries are the published
                    // column (other label cj / ci) or the stashed row (own label ci / cj).
#pragma unroll
                    for (int m = 0; m < NBT; ++m) {
                        const int b = tid + m * NT;
                        if (b < NB) {
                            const int ea = bend[m] & 0xFFFF, et = bend[m] >> 16;
                            const int wdn = (1 << 16) | cj_j;
                            const int ia = ea == sa_j ? wdn : (ea == sb_j ? ci_j : binfo[m][0]);
                            const int it = et == sa_j ? wdn : (et == sb_j ? ci_j : binfo[m][1]);
                            binfo[m][0] = ia;
                            binfo[m][1] = it;
                            // every operand is loaded in one batch (each label indexes valid
                                // rows also on a parallel bond); only antiparallel bonds store
                            {
                                const bool aup = (ia >> 16) == 0;
                                const int i = (aup ? ia : it) & 0xFFFF, jd = (aup ? it : ia) & 0xFFFF;
                                // G_up[ci][jd] + delta(jd, cj) and G_dn[cj][i] + delta(i, ci)
                                const C v0 = MF ? cmk<C>(mrow[jd], mrow[N + jd])
                                                : cmk<C>(s_row[0][jd % LPR][0][jd / LPR], s_row[0][jd % LPR][1][jd / LPR]);
                                const C v1 = MF ? cmk<C>(mrow[2 * N + i], mrow[3 * N + i])
                                                : cmk<C>(s_row[1][i % LPR][0][i / LPR], s_row[1][i % LPR][1][i / LPR]);
                                const bool oi = i == ci_j, oj = jd == cj_j;
                                const C c0 = s_bq[b][0], c1 = s_bq[b][1];
                                const C x0 = MF ? cmk<C>(mux[i], mux[N + i]) : s_ux[0][i];
                                const C x1 = MF ? cmk<C>(mux[2 * N + jd], mux[3 * N + jd]) : s_ux[1][jd];
                                const C n0 = MF ? mf_coefficient(0, i) : s_nc[0][i];
                                const C n1 = MF ? mf_coefficient(1, jd) : s_nc[1][jd];
                                const C o0 = (oj && !oi) ? x0 : ((oi && !oj) ? v0 : c0);
                                const C o1 = (oi && !oj) ? x1 : ((oj && !oi) ? v1 : c1);
                                C r0, r1;
                                r0.x = fma(n0.x, v0.x, o0.x);
                                r0.x = fma(-n0.y, v0.y, r0.x);
                                r0.y = fma(n0.x, v0.y, o0.y);
                                r0.y = fma(n0.y, v0.x, r0.y);
                                r1.x = fma(n1.x, v1.x, o1.x);
                                r1.x = fma(-n1.y, v1.y, r1.x);
                                r1.y = fma(n1.x, v1.y, o1.y);
                                r1.y = fma(n1.y, v1.x, r1.y);
                                // (a parallel bond keeps its entries: no branch around the store)
                                const bool apb = ((ia ^ it) >> 16) != 0;
                                s_bq[b][0] = apb ? r0 : c0;
                                s_bq[b][1] = apb ? r1 : c1;
                            }
                        }
                    }
                };
                // matrix-core tables: the bond update first, in the same block as the MFMAs, so
                // that its loads and arithmetic issue while the matrix pipe works through them
                if constexpr (MF && BQ) bond_update();
                if constexpr (MF) {
#pragma unroll
                    for (int fl = 0; fl < 2; ++fl) {
                        const C nc = mf_coefficient(fl, mrb * 32 + ml);
                        const float ar = mh ? -nc.y : nc.x, ai = mh ? nc.x : nc.y;
#pragma unroll
                        for (int cb = 0; cb < 2; ++cb) {
                            const float bv = mrow[(fl * 2 + mh) * N + (2 * mcp + cb) * 32 + ml];
                            mr[fl][cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, bv, mr[fl][cb], 0, 0, 0);
                            mi[fl][cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(ai, bv, mi[fl][cb], 0, 0, 0);
                        }
                    }
                }
#pragma unroll
                for (int fl = 0; fl < 2; ++fl) {
                    C uu;
                    const C nc = MF ? cmk<C>(0, 0) : (BQ ? ncb[fl] : coefficient(fl, uu));
                    if (!MF && has_row) {
                        const T *sr = &s_row[fl][c][0][0], *si = &s_row[fl][c][1][0];
                        if constexpr (RES) {
#pragma unroll
                            for (int k = 0; k < CPT; ++k) {
                                const T vr = sr[k], vi = si[k];
                                T er = gr[fl][k], ei = gi[fl][k];
                                er = fma(nc.x, vr, er);
                                er = fma(-nc.y, vi, er);
                                ei = fma(nc.x, vi, ei);
                                ei = fma(nc.y, vr, ei);
                                gr[fl][k] = er;
                                gi[fl][k] = ei;
                            }
                        } else {
                            // streamed rows in chunks of SK slots, fully unrolled: the next
                            // chunk's loads are issued before this chunk's stores, so that the
                            // loads do not wait behind the stores
                            constexpr int SK = stream_chunk(CPT), NCH = CPT / SK;
                            C ea[SK], eb[SK];
                            auto ldk = [&](int k0, C (&e)[SK]) {
#pragma unroll
                                for (int j = 0; j < SK; ++j) e[j] = gat(fl, row, c + LPR * (k0 + j));
                            };
                            auto upk = [&](int k0, C (&e)[SK]) {
#pragma unroll
                                for (int j = 0; j < SK; ++j) {
                                    const int k = k0 + j;
                                    gat(fl, row, c + LPR * k) = cfma(nc, cmk<C>(sr[k], si[k]), e[j]);
                                }
                            };
                            ldk(0, ea);
#pragma unroll
                            for (int ch = 0; ch < NCH; ch += 2) {
                                if (ch + 1 < NCH) ldk((ch + 1) * SK, eb);
                                upk(ch * SK, ea);
                                if (ch + 2 < NCH) ldk((ch + 2) * SK, ea);
                                if (ch + 1 < NCH) upk((ch + 1) * SK, eb);
                            }
                        }
                    }
                }
                if constexpr (BQ) {
                    if constexpr (!MF) bond_update();
                } else {
                    // every wavefront writes the same values, so each one reads back its own
                    // writes in the next window
                    if (lane == 0) hop();
                    if (tid == 0) amplitude();
                }
                // streamed tables: every row is rewritten before the next window reads them;
                // BQ: the bond cache is current before the next window reads it
                if constexpr (!RES || BQ) __syncthreads();
            }
            s0 += nproc;
            par ^= 1;
            const int last = s0 - 1;
            if (meas > 0 && last == next_meas) {
                measure(last);
                next_meas = last + meas;
            }
        }
    } else {
        // Delayed-update windows.  A window is up to WD_K steps (inside the batch, ending
        // at the next measured step) with at most WD_M accepted moves.  The particles the
        // window can touch are those at its sites: the up labels U and down labels D found
        // there (a swap only exchanges particles between the window's sites), so every
        // ratio the window needs is an entry of the sub-blocks G_up[U][D] and G_dn[D][U].
        //  1. the owners of the rows U of G_up and D of G_dn copy them to an LDS dump
        //     (double-buffered by window parity); barrier;
        //  2. wavefront 0 runs MetroMC::Step for each step in order.  An entry after n moves
        //     is the dumped entry with the n rank-1 terms of the moves applied in order, the
        //     table's own formula; move n stores its factors x_n (column - e) and y_n
        //     (row + e) by sub-block position.  From them follows the m x m matrix M with
        //        G_after = G - C M R,   C[:,b] = G[:,col_b] - e_row_b,  R[d,:] = G[row_d,:] + e_col_d
        //     (x_n = sum_b S_nb C[:,b], y_n = sum_d T_nd R[d,:], S and T unit lower
        //     triangular from gamma_nm = y_m[col_n] / q_m and beta_nm = x_m[row_n] / q_m,
        //     M = S^T diag(1/q) T); it writes the moves, counts and traces; barrier;
        //  3. every thread applies the rank-m update to its rows, R from the dump.
        // The decisions are MetroMC::Step's on the same uniforms in the same order; the
        // tables after the window equal the m rank-1 updates up to rounding.
        constexpr int NMW = (N + 63) / 64;
        constexpr size_t DE = wdu_dump_elems<T, N, LPR>();
        static_assert(NT / 64 <= 16, "mask slots");
        WduScratch<T> &ws = *reinterpret_cast<WduScratch<T> *>(smem + 2 * DE * sizeof(T));
        const int wu = __builtin_amdgcn_readfirstlane(wid);  // wave-uniform: scalar branches
        int bpv = 0;
        double bu3 = 1.0;
        int par = 0;
        int s0 = 0;
        int next_meas = 0;  // the next measured step (meas > 0)
        while (s0 < n_steps) {
            const int b0 = (int)(s0 & 63);
            if (b0 == 0) draw_batch(s0, bpv, bu3);
            int lim = n_steps - s0;
            if (lim > 64 - b0) lim = 64 - b0;
            if (meas > 0 && next_meas - s0 + 1 < lim) lim = next_meas - s0 + 1;
            const int len = (int)(lim < WD_K ? lim : WD_K);
            const bool inw = lane >= b0 && lane < b0 + len;
            const int v1 = bpv & 0xFFFF, v2 = bpv >> 16;
            const double u3 = bu3;
            const int i1 = s_site[v1], i2 = s_site[v2];
            // label masks of the window's sites: LDS atomic ORs into this wavefront's slot
            // (every wavefront derives the same masks)
            uint32_t *mk = &ws.msk[wu][0][0];
            if (lane < 8) mk[lane] = 0u;
            wave_lds_order();
            if (inw) {
                atomicOr(&ws.msk[wu][i1 >> 16][(i1 & 0xFFFF) >> 5], 1u << (i1 & 31));
                atomicOr(&ws.msk[wu][i2 >> 16][(i2 & 0xFFFF) >> 5], 1u << (i2 & 31));
            }
            wave_lds_order();
            uint64_t um[NMW], dm[NMW];
#pragma unroll
            for (int q = 0; q < NMW; ++q) {
                um[q] = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)mk[2 * q]) |
                        ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)mk[2 * q + 1]) << 32);
                dm[q] = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)mk[4 + 2 * q]) |
                        ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)mk[4 + 2 * q + 1]) << 32);
            }
            int nU = 0, nD = 0;
#pragma unroll
            for (int q = 0; q < NMW; ++q) {
                nU += __popcll(um[q]);
                nD += __popcll(dm[q]);
            }
            T *dump = reinterpret_cast<T *>(smem) + (size_t)par * DE;
            // element (dump row p, column label l)
            auto dat = [&](int p, int l) -> C {
                const T *e = dump + ((size_t)(p * LPR + l % LPR) * 2) * RP + l / LPR;
                return cmk<C>(e[0], e[RP]);
            };
            if (has_row) {
                const int rw = row >> 6;
                uint64_t ur = um[0], dr = dm[0];
                if constexpr (NMW > 1) {
                    ur = rw ? um[1] : um[0];
                    dr = rw ? dm[1] : dm[0];
                }
#pragma unroll
                for (int fl = 0; fl < 2; ++fl) {
                    if (((fl ? dr : ur) >> (row & 63)) & 1ull) {
                        const int p = fl ? nU + mask_rank<NMW>(dm, row) : mask_rank<NMW>(um, row);
                        T *d = dump + ((size_t)(p * LPR + c) * 2) * RP;
#pragma unroll
                        for (int k = 0; k < CPT; ++k) {
                            d[k] = gr[fl][k];
                            d[RP + k] = gi[fl][k];
                        }
                    }
                }
            }
            if (wu == 0 && inw) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int x = h ? i2 : i1, l = x & 0xFFFF;
                    if (x >> 16) ws.lab[1][mask_rank<NMW>(dm, l)] = l;
                    else ws.lab[0][mask_rank<NMW>(um, l)] = l;
                }
            }
            __syncthreads();
            if (wu == 0) {
                // entry (row position r, column position cp) of table tb after the first n
                // moves, from the dumped entry v: the moves' rank-1 terms in order
                auto moved_entry = [&](C v, int n, int tb, int r, int cp) -> C {
#pragma unroll
                    for (int mm = 0; mm < WD_M; ++mm) {
                        if (mm < n) {
                            const C cf = ws.vec[mm][2 * tb][r], y = ws.vec[mm][2 * tb + 1][cp];
                            T er = v.x, ei = v.y;
                            er = fma(-cf.x, y.x, er);
                            er = fma(cf.y, y.y, er);
                            ei = fma(-cf.x, y.y, ei);
                            ei = fma(-cf.y, y.x, ei);
                            v = cmk<C>(er, ei);
                        }
                    }
                    return v;
                };
                const int lu = lane < nU ? ws.lab[0][lane] : 0;  // label of U position `lane`
                const int ld = lane < nD ? ws.lab[1][lane] : 0;  // label of D position `lane`
                // the particles at the step's two sites as the window's moves leave them,
                // (flavour << 16) | label
                int x1 = i1, x2 = i2;
                int m = 0, sst = b0, jend = b0 + len - 1;
                uint64_t noop_m = 0ull;  // finished steps that were trivial
                for (;;) {
                    const bool act = inw && lane >= sst;
                    const bool noop = ((x1 ^ x2) >> 16) == 0;
                    const bool up1 = (x1 >> 16) == 0;
                    const int ci = (up1 ? x1 : x2) & 0xFFFF, cj = (up1 ? x2 : x1) & 0xFFFF;
                    int pa = 0, pb = 0;
                    C qu = cmk<C>(0, 0), qd = cmk<C>(0, 0);
                    if (act && !noop) {
                        pa = mask_rank<NMW>(um, ci);
                        pb = mask_rank<NMW>(dm, cj);
                        qu = moved_entry(dat(pa, cj), m, 0, pa, pb);
                        qd = moved_entry(dat(nU + pb, ci), m, 1, pb, pa);
                    }
                    double ratio = 1.0;
                    bool accl;
                    if (noop) {
                        accl = u3 < 1.0;  // LatticeStepper.cpp:64-68: ratio 1
                    } else {
                        const double n0 = (double)qu.x * qu.x + (double)qu.y * qu.y;
                        const double n1 = (double)qd.x * qd.x + (double)qd.y * qd.y;
                        ratio = n0 * n1;
                        accl = ratio > 1.0 ? (u3 < 1.0) : (u3 < ratio);  // MetroMC::YesNo
                    }
                    const uint64_t moved = __ballot(act && !noop && accl);
                    const int j = moved ? (int)__builtin_ctzll(moved) : 64;
                    // steps sst .. j are finished with these decisions
                    const bool fin = act && lane <= j;
                    noop_m |= __ballot(fin && noop);
                    if (fin) {
                        const size_t ti = (size_t)w * n_steps + s0 + (lane - b0);
                        if (tr_noop) tr_noop[ti] = noop ? 1 : 0;
                        if (tr_accept) tr_accept[ti] = accl ? 1 : 0;
                        if (tr_ratio) tr_ratio[ti] = ratio;
                    }
                    if (!moved) break;
                    const int ci_j = __builtin_amdgcn_readlane(ci, j), cj_j = __builtin_amdgcn_readlane(cj, j);
                    const int pa_j = __builtin_amdgcn_readlane(pa, j), pb_j = __builtin_amdgcn_readlane(pb, j);
                    const int u1_j = __builtin_amdgcn_readlane(up1 ? 1 : 0, j);
                    const int v1_j = __builtin_amdgcn_readlane(v1, j), v2_j = __builtin_amdgcn_readlane(v2, j);
                    const int sa_j = u1_j ? v1_j : v2_j, sb_j = u1_j ? v2_j : v1_j;
                    const C qU = cmk<C>(readlane_real(qu.x, j), readlane_real(qu.y, j));
                    const C qD = cmk<C>(readlane_real(qd.x, j), readlane_real(qd.y, j));
                    // the move's factors: up table column cj (over U) and row ci (over D),
                    // down table column ci (over D) and row cj (over U), after the earlier moves
                    C xu_ = cmk<C>(0, 0), yd_ = cmk<C>(0, 0), yu_ = cmk<C>(0, 0), xd_ = cmk<C>(0, 0);
                    if (lane < nU) {
                        xu_ = moved_entry(dat(lane, cj_j), m, 0, lane, pb_j);
                        yd_ = moved_entry(dat(nU + pb_j, lu), m, 1, pb_j, lane);
                    }
                    if (lane < nD) {
                        yu_ = moved_entry(dat(pa_j, ld), m, 0, pa_j, lane);
                        xd_ = moved_entry(dat(nU + lane, ci_j), m, 1, lane, pa_j);
                    }
                    // gamma and beta of this move against the earlier ones (FP64)
                    if (lane < 4 * m) {
                        const int mb = lane >> 2, wh = lane & 3, tb = wh >> 1, gb = wh & 1;
                        // gamma: y_m at this move's column; beta: x_m at this move's row
                        const int off = tb == 0 ? (gb ? pa_j : pb_j) : (gb ? pb_j : pa_j);
                        const C v = gb ? ws.xr[mb][tb][off] : ws.vec[mb][2 * tb + 1][off];
                        const double2 iq = ws.iq[tb][mb];
                        ws.coef[tb][gb][m][mb] = make_double2((double)v.x * iq.x - (double)v.y * iq.y,
                                                              (double)v.x * iq.y + (double)v.y * iq.x);
                    }
                    {
                        // 1/q as the table update forms it (FP64 norm, FP32 factor)
                        const double nu = (double)qU.x * qU.x + (double)qU.y * qU.y;
                        const double nd = (double)qD.x * qD.x + (double)qD.y * qD.y;
                        const C qiu = cmk<C>(qU.x / nu, -qU.y / nu), qid = cmk<C>(qD.x / nd, -qD.y / nd);
                        if (lane < nU) {
                            if (lane == pa_j) {
                                xu_.x -= (T)1;
                                yd_.x += (T)1;
                            }
                            ws.xr[m][0][lane] = xu_;
                            ws.vec[m][0][lane] = cmul(xu_, qiu);
                            ws.vec[m][3][lane] = yd_;
                        }
                        if (lane < nD) {
                            if (lane == pb_j) {
                                yu_.x += (T)1;
                                xd_.x -= (T)1;
                            }
                            ws.vec[m][1][lane] = yu_;
                            ws.xr[m][1][lane] = xd_;
                            ws.vec[m][2][lane] = cmul(xd_, qid);
                        }
                        if (lane < 2) {
                            const C q = lane ? qD : qU;
                            const double nq = lane ? nd : nu;
                            ws.iq[lane][m] = make_double2((double)q.x / nq, -(double)q.y / nq);
                            ws.qd[lane][m] = make_double2((double)q.x, (double)q.y);
                        }
                    }
                    if (lane == 0) {
                        ws.mv[m][0] = ci_j;
                        ws.mv[m][1] = cj_j;
                        ws.mv[m][2] = sa_j;
                        ws.mv[m][3] = sb_j;
                        ws.pos[m][0] = pa_j;
                        ws.pos[m][1] = pb_j;
                    }
                    wave_lds_order();
                    // State::Hop on the window's sites
                    const int xa = (1 << 16) | cj_j, xb = ci_j;
                    x1 = v1 == sa_j ? xa : (v1 == sb_j ? xb : x1);
                    x2 = v2 == sa_j ? xa : (v2 == sb_j ? xb : x2);
                    ++m;
                    sst = j + 1;
                    if (m == WD_M || sst >= b0 + len) {
                        jend = j;
                        break;
                    }
                }
                if (m > 0) {
                    // columns of S (tb, 0) and T (tb, 1): v_n = delta_nd - sum_b coef_nb v_b
                    if (lane < 4 * WD_M) {
                        const int tb = lane / (2 * WD_M), gb = (lane / WD_M) & 1, d = lane % WD_M;
                        double2 v[WD_M];
#pragma unroll
                        for (int n = 0; n < WD_M; ++n) {
                            double2 t = make_double2(n == d ? 1.0 : 0.0, 0.0);
                            if (n < m) {
#pragma unroll
                                for (int b = 0; b < n; ++b) {
                                    const double2 cf = ws.coef[tb][gb][n][b];
                                    t.x -= cf.x * v[b].x - cf.y * v[b].y;
                                    t.y -= cf.x * v[b].y + cf.y * v[b].x;
                                }
                                ws.st[tb][gb][n][d] = t;
                            }
                            v[n] = t;
                        }
                    }
                    wave_lds_order();
                    if (lane < 2 * WD_M * WD_M) {
                        const int tb = lane / (WD_M * WD_M), b = (lane / WD_M) % WD_M, d = lane % WD_M;
                        if (b < m && d < m) {
                            double2 acc = make_double2(0.0, 0.0);
                            for (int n = 0; n < m; ++n) {
                                const double2 sv = ws.st[tb][0][n][b], tv = ws.st[tb][1][n][d], iq = ws.iq[tb][n];
                                const double pr = sv.x * tv.x - sv.y * tv.y, pi = sv.x * tv.y + sv.y * tv.x;
                                acc.x += pr * iq.x - pi * iq.y;
                                acc.y += pr * iq.y + pi * iq.x;
                            }
                            ws.mm[tb][b][d] = cmk<C>(acc.x, acc.y);
                        }
                    }
                }
                const int nproc = jend - b0 + 1;
                const int nnoop = __popcll(noop_m);
                n_noop += nnoop;
                n_acc += m;
                n_rej += nproc - nnoop - m;
                if (lane == 0) {
                    ws.res[0] = m;
                    ws.res[1] = nproc;
                    // State::Hop (State.cpp:73-84) of each move, in order
                    for (int n = 0; n < m; ++n) {
                        const int ci_n = ws.mv[n][0], cj_n = ws.mv[n][1], sa_n = ws.mv[n][2], sb_n = ws.mv[n][3];
                        s_site[sb_n] = ci_n;
                        s_site[sa_n] = (1 << 16) | cj_n;
                        s_part[0][ci_n] = sb_n;
                        s_part[1][cj_n] = sa_n;
                    }
                }
                if (lane < 2 && m > 0) {
                    // amplitude *= the moves' ratios of flavour `lane` (SlaterDeterminant.cpp:321)
                    const int fl = lane;
                    double mr = s_amp[2 * fl], mi = s_amp[2 * fl + 1], ex = s_amp[4 + fl];
                    for (int n = 0; n < m; ++n) {
                        const double2 q = ws.qd[fl][n];
                        amp_mul(mr, mi, ex, q.x, q.y);
                    }
                    s_amp[2 * fl] = mr;
                    s_amp[2 * fl + 1] = mi;
                    s_amp[4 + fl] = ex;
                }
            }
            __syncthreads();
            const int m = __builtin_amdgcn_readfirstlane(ws.res[0]);
            const int nproc = __builtin_amdgcn_readfirstlane(ws.res[1]);
            if (m > 0 && has_row) {
#pragma unroll
                for (int fl = 0; fl < 2; ++fl) {
                    int colv[WD_M], rowv[WD_M], dpos[WD_M];
                    C wv[WD_M];
#pragma unroll
                    for (int d = 0; d < WD_M; ++d) wv[d] = cmk<C>(0, 0);
#pragma unroll
                    for (int b = 0; b < WD_M; ++b) {
                        colv[b] = rowv[b] = dpos[b] = 0;
                        if (b < m) {
                            const int ci_b = __builtin_amdgcn_readfirstlane(ws.mv[b][0]);
                            const int cj_b = __builtin_amdgcn_readfirstlane(ws.mv[b][1]);
                            colv[b] = fl ? ci_b : cj_b;
                            rowv[b] = fl ? cj_b : ci_b;
                            dpos[b] = fl ? nU + __builtin_amdgcn_readfirstlane(ws.pos[b][1])
                                         : __builtin_amdgcn_readfirstlane(ws.pos[b][0]);
                            // C[row][b] = G[row][col_b] - delta(row, row_b), from lane col_b % LPR
                            const int k = colv[b] / LPR, ls = colv[b] % LPR;
                            T xr = gr[fl][k], xi = gi[fl][k];
                            xr = __shfl(xr, gbase + ls);
                            xi = __shfl(xi, gbase + ls);
                            if (row == rowv[b]) xr -= (T)1;
                            // w_d += C[row][b] M[b][d]
                            const C cbv = cmk<C>(xr, xi);
#pragma unroll
                            for (int d = 0; d < WD_M; ++d)
                                if (d < m) wv[d] = cfma(cbv, ws.mm[fl][b][d], wv[d]);
                        }
                    }
#pragma unroll
                    for (int d = 0; d < WD_M; ++d) {
                        if (d < m) {
                            const C wd = wv[d];
                            // G[row][:] -= w (G[row_d][:] + e_col_d)
                            const T *sr = dump + ((size_t)(dpos[d] * LPR + c) * 2) * RP, *si = sr + RP;
#pragma unroll
                            for (int k = 0; k < CPT; ++k) {
                                const T vr = sr[k], vi = si[k];
                                T er = gr[fl][k], ei = gi[fl][k];
                                er = fma(-wd.x, vr, er);
                                er = fma(wd.y, vi, er);
                                ei = fma(-wd.x, vi, ei);
                                ei = fma(-wd.y, vr, ei);
                                gr[fl][k] = er;
                                gi[fl][k] = ei;
                            }
                            if (c == colv[d] % LPR) {
                                const int k = colv[d] / LPR;
                                gr[fl][k] -= wd.x;
                                gi[fl][k] -= wd.y;
                            }
                        }
                    }
                }
            }
            s0 += nproc;
            par ^= 1;
            const int last = s0 - 1;
            if (meas > 0 && last == next_meas) {
                measure(last);
                next_meas = last + meas;
            }
        }
    }

    flush_snaps();
    __syncthreads();
    if constexpr (MF) {
        // the element addresses are formed anew here (an opaque base), so none of the load's
        // addresses stays live through the walk
        C *gS = gG;
        asm volatile("" : "+s"(gS));
#pragma unroll
        for (int fl = 0; fl < 2; ++fl)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    gS[(size_t)fl * N * N + (size_t)(mrb * 32 + (r >> 2) * 8 + mh * 4 + (r & 3)) * N +
                       (2 * mcp + cb) * 32 + ml] = cmk<C>(mr[fl][cb][r], mi[fl][cb][r]);
    } else if constexpr (RES && !MONLY) {
        if (has_row) {
#pragma unroll
            for (int fl = 0; fl < 2; ++fl)
#pragma unroll UNR
                for (int k = 0; k < CPT; ++k) gat(fl, row, c + LPR * k) = cmk<C>(gr[fl][k], gi[fl][k]);
        }
    }
    for (int i = tid; i < NS; i += NT) a.site[(size_t)w * NS + i] = s_site[i];
    for (int i = tid; i < 2 * N; i += NT) a.part[(size_t)w * 2 * N + i] = s_part[i / N][i % N];
    if (tid == 0) {
#pragma unroll
        for (int fl = 0; fl < 2; ++fl) {
            const double mr = s_amp[2 * fl], mi = s_amp[2 * fl + 1];
            a.logabs[2 * w + fl] += log(hypot(mr, mi)) + s_amp[4 + fl] * 0.69314718055994531;
            a.phase[2 * w + fl] = remainder(a.phase[2 * w + fl] + atan2(mi, mr), 2 * M_PI);
        }
        a.counter[w] = ctr0 + (uint64_t)a.n_steps;
        s_acc[7] = MONLY ? 0.0 : (double)a.n_steps;
        s_acc[8] = (double)n_acc;
        s_acc[9] = (double)n_rej;
        s_acc[10] = (double)n_noop;
        double *ac = a.acc + (size_t)w * ACC_STRIDE;
        for (int i = 0; i < ACC_STRIDE; ++i) ac[i] += s_acc[i];
    }
}

// ---------------------------------------------------------------------------------
// Ratio tables from the FP64 inverses, one block per (walker, flavour):
// G_fl[i][j] = sum_f Ai_fl[i][f] Phi_fl(orb_f, site of particle j of the other flavour),
// accumulated in FP64 over LDS-staged chunks of FB orbitals and written as OUT.  If
// `drift` is given, max|G_old - G_new| / max|G_new| is stored first (drift refresh).
// The thread's CPT output slots are formed CB at a time, CB sized to the VGPR budget of
// the launch shape (one pass unless N is large).
template <typename OUT, int N, int LPR>
__global__ void __launch_bounds__(((N * LPR + 63) / 64) * 64)
gbuild_kernel(GBArgs a)
{
    using C = double2;
    using CO = typename cpx<OUT>::t;
    constexpr int NT = ((N * LPR + 63) / 64) * 64;
    constexpr int NW = NT / 64;
    constexpr int CPT = N / LPR;
    constexpr int NS = 2 * N;
    constexpr int FB = (N % 8 == 0) ? 8 : ((N % 6 == 0) ? 6 : ((N % 2 == 0) ? 2 : 1));
    static_assert(N % FB == 0, "orbital chunk must divide N");
    constexpr int WPS = (NW + 3) / 4;                             // waves per SIMD
    constexpr int VB = (512 / WPS < 256 ? 512 / WPS : 256) / 2;   // VGPRs for accumulators
    constexpr int CB = divisor_within(CPT, 4, VB);                // slots per pass

    __shared__ int s_other[N];        // sites of the other flavour's particles
    __shared__ C s_ai[N][FB + 1];     // +1 pads the row stride against bank conflicts
    __shared__ C s_b[FB][N];
    __shared__ double s_red[NW][2];

    const int w = blockIdx.x >> 1, fl = blockIdx.x & 1;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int row = tid / LPR, c = tid % LPR;
    const bool has_row = row < N;
    for (int i = tid; i < N; i += NT) s_other[i] = a.part[(size_t)w * 2 * N + (size_t)(1 - fl) * N + i];
    const C *ai = a.Ai + (size_t)w * 2 * N * N + (size_t)fl * N * N;
    const C *occ = a.occ + (size_t)a.group[w] * 2 * NS * N + (size_t)fl * NS * N;
    CO *out = reinterpret_cast<CO *>(a.G) + (size_t)w * 2 * N * N + (size_t)fl * N * N;
    double dmax = 0.0, nmax = 0.0;
    __syncthreads();
    for (int k0 = 0; k0 < CPT; k0 += CB) {
        C acc[CB];
#pragma unroll
        for (int k = 0; k < CB; ++k) acc[k] = make_double2(0.0, 0.0);
        for (int f0 = 0; f0 < N; f0 += FB) {
            for (int i = tid; i < N * FB; i += NT) {
                const int r = i / FB, ff = i % FB;
                s_ai[r][ff] = ai[(size_t)r * N + f0 + ff];
                const int j = i / FB;
                s_b[ff][j] = occ[(size_t)s_other[j] * N + f0 + ff];
            }
            __syncthreads();
            if (has_row) {
#pragma unroll
                for (int ff = 0; ff < FB; ++ff) {
                    const C x = s_ai[row][ff];
#pragma unroll
                    for (int k = 0; k < CB; ++k) acc[k] = cfma(x, s_b[ff][c + LPR * (k0 + k)], acc[k]);
                }
            }
            __syncthreads();
        }
        if (has_row) {
#pragma unroll
            for (int k = 0; k < CB; ++k) {
                const size_t o = (size_t)row * N + c + LPR * (k0 + k);
                if (a.drift) {
                    const CO old = out[o];
                    const double dx = (double)old.x - acc[k].x, dy = (double)old.y - acc[k].y;
                    dmax = fmax(dmax, sqrt(dx * dx + dy * dy));
                    nmax = fmax(nmax, sqrt(acc[k].x * acc[k].x + acc[k].y * acc[k].y));
                }
                CO v;
                v.x = (OUT)acc[k].x;
                v.y = (OUT)acc[k].y;
                out[o] = v;
            }
        }
    }
    if (a.drift) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            dmax = fmax(dmax, __shfl_xor(dmax, o));
            nmax = fmax(nmax, __shfl_xor(nmax, o));
        }
        if (lane == 0) { s_red[wid][0] = dmax; s_red[wid][1] = nmax; }
        __syncthreads();
        if (tid == 0) {
            double d = 0.0, n = 0.0;
            for (int q = 0; q < NW; ++q) { d = fmax(d, s_red[q][0]); n = fmax(n, s_red[q][1]); }
            a.drift[2 * w + fl] = n > 0.0 ? d / n : 0.0;
        }
    }
}

// The same tables on the FP64 matrix cores (v_mfma_f64_16x16x4f64), N a multiple of 16:
// wave w forms the 16-row band [16 w, 16 w + 16) of G over all N / 16 column tiles, the
// complex product as four real products per tile (re += Ar Br - Ai Bi, im += Ar Bi + Ai Br).
// Orbitals go through LDS in chunks of KC; the next chunk's global loads are in flight
// while the current one is multiplied.  Operand layouts: A lane l = row l % 16, k l / 16;
// B lane l = k l / 16, column l % 16; result element v of lane l = row l / 16 + 4 v,
// column l % 16.
template <typename OUT, int N>
__global__ void __launch_bounds__(N * 4) gbuild_mfma_kernel(GBArgs a)
{
    using C = double2;
    using CO = typename cpx<OUT>::t;
    typedef double v4 __attribute__((ext_vector_type(4)));
    constexpr int NW = N / 16, NT = 64 * NW, NCB = N / 16;
    constexpr int KC = 16;                       // orbitals per LDS chunk
    constexpr int PER = N * KC / NT;             // chunk elements per thread, per operand
    static_assert(N % 16 == 0 && N % KC == 0 && (N * KC) % NT == 0, "N: a multiple of 16");

    __shared__ int s_other[N];
    __shared__ C s_a[N][KC + 1];                 // Ai chunk, [row][orbital]
    __shared__ C s_bm[KC][N + 1];                // Phi chunk, [orbital][particle]
    __shared__ double s_red[NW][2];

    const int w = blockIdx.x >> 1, fl = blockIdx.x & 1;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int li = lane & 15, kq = lane >> 4;
    for (int i = tid; i < N; i += NT) s_other[i] = a.part[(size_t)w * 2 * N + (size_t)(1 - fl) * N + i];
    const C *ai = a.Ai + (size_t)w * 2 * N * N + (size_t)fl * N * N;
    const C *occ = a.occ + (size_t)a.group[w] * 2 * 2 * N * N + (size_t)fl * 2 * N * N;
    CO *out = reinterpret_cast<CO *>(a.G) + (size_t)w * 2 * N * N + (size_t)fl * N * N;
    __syncthreads();

    // chunk element e = tid + NT i of each operand: row (or particle) e / KC, orbital e % KC
    // (plain doubles: the vector type would keep the in-flight chunk in scratch)
    double par[PER], pai[PER], pbr[PER], pbi[PER];
    auto fetch = [&](int f0) __attribute__((always_inline)) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = tid + NT * i, r = e / KC, ff = e % KC;
            const C x = ai[(size_t)r * N + f0 + ff];
            const C y = occ[(size_t)s_other[r] * N + f0 + ff];
            par[i] = x.x;
            pai[i] = x.y;
            pbr[i] = y.x;
            pbi[i] = y.y;
        }
    };
    fetch(0);
    v4 cr[NCB], cim[NCB];
#pragma unroll
    for (int t = 0; t < NCB; ++t) {
        cr[t] = v4{0.0, 0.0, 0.0, 0.0};
        cim[t] = v4{0.0, 0.0, 0.0, 0.0};
    }
#pragma unroll 1
    for (int f0 = 0; f0 < N; f0 += KC) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = tid + NT * i, r = e / KC, ff = e % KC;
            s_a[r][ff] = make_double2(par[i], pai[i]);
            s_bm[ff][r] = make_double2(pbr[i], pbi[i]);
        }
        __syncthreads();
        if (f0 + KC < N) fetch(f0 + KC);
#pragma unroll
        for (int kk = 0; kk < KC / 4; ++kk) {
            const C av = s_a[wid * 16 + li][kk * 4 + kq];
            const double nai = -av.y;
            C bv[NCB];
#pragma unroll
            for (int t = 0; t < NCB; ++t) bv[t] = s_bm[kk * 4 + kq][t * 16 + li];
#pragma unroll
            for (int t = 0; t < NCB; ++t) cr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv[t].x, cr[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < NCB; ++t) cim[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv[t].y, cim[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < NCB; ++t) cr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, bv[t].y, cr[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < NCB; ++t) cim[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, bv[t].x, cim[t], 0, 0, 0);
        }
    }
    double dmax = 0.0, nmax = 0.0;
#pragma unroll
    for (int t = 0; t < NCB; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = wid * 16 + kq + 4 * v, col = t * 16 + li;
            const size_t o = (size_t)row * N + col;
            const double gx = cr[t][v], gy = cim[t][v];
            if (a.drift) {
                const CO old = out[o];
                const double dx = (double)old.x - gx, dy = (double)old.y - gy;
                dmax = fmax(dmax, sqrt(dx * dx + dy * dy));
                nmax = fmax(nmax, sqrt(gx * gx + gy * gy));
            }
            CO g;
            g.x = (OUT)gx;
            g.y = (OUT)gy;
            out[o] = g;
        }
    if (a.drift) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            dmax = fmax(dmax, __shfl_xor(dmax, o));
            nmax = fmax(nmax, __shfl_xor(nmax, o));
        }
        if (lane == 0) { s_red[wid][0] = dmax; s_red[wid][1] = nmax; }
        __syncthreads();
        if (tid == 0) {
            double d = 0.0, n = 0.0;
            for (int q = 0; q < NW; ++q) { d = fmax(d, s_red[q][0]); n = fmax(n, s_red[q][1]); }
            a.drift[2 * w + fl] = n > 0.0 ? d / n : 0.0;
        }
    }
}

// ---------------------------------------------------------------------------------
// Sum of per-walker StatSpinStruct blocks over the walkers of each parameter group:
// blockIdx.y = group, one thread per entry, walkers read in order (coalesced over e).
__global__ void __launch_bounds__(256) ss_reduce_kernel(const double *ss, const int32_t *idx,
                                                        const int32_t *off, size_t per, double *out)
{
    const int g = blockIdx.y;
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= per) return;
    double t = 0.0;
    for (int k = off[g]; k < off[g + 1]; ++k) t += ss[(size_t)idx[k] * per + e];
    out[(size_t)g * per + e] = t;
}

// ---------------------------------------------------------------------------------
// host-side dispatch tables
template <typename T, int N, int LPR, bool RES>
static hipError_t launch_sweep_t(const SweepArgs &a, int nblocks, hipStream_t st)
{
    constexpr int NT = ((N * LPR + 63) / 64) * 64;
    // the measurement-only pass (gpvmc_measure) is its own kernel symbol, so profiles of
    // the sweep report sweeps only
    const bool traced = a.uniforms || a.tr_noop || a.tr_accept || a.tr_ratio || a.tr_eloc || a.tr_mz ||
                        a.tr_logw;
    // FP32 tables of N = 128 on the matrix cores, except for StatSpinStruct, which reads the
    // tables in their row layout
    constexpr bool MFOK = RES && N == 128 && LPR == 4 && sizeof(T) == 4;
    if (MFOK && !a.measure_only && !(a.mask & MEAS_STATSPIN)) {
        if (traced)
            hipLaunchKernelGGL((sweep_kernel<T, N, LPR, RES, false, false, true, MFOK>), dim3(nblocks), dim3(NT), 0,
                               st, a);
        else
            hipLaunchKernelGGL((sweep_kernel<T, N, LPR, RES, false, false, false, MFOK>), dim3(nblocks), dim3(NT), 0,
                               st, a);
    } else if (a.measure_only)
        hipLaunchKernelGGL((sweep_kernel<T, N, LPR, RES, true>), dim3(nblocks), dim3(NT), 0, st, a);
    else if (traced)
        hipLaunchKernelGGL((sweep_kernel<T, N, LPR, RES, false>), dim3(nblocks), dim3(NT), 0, st, a);
    else
        hipLaunchKernelGGL((sweep_kernel<T, N, LPR, RES, false, false, false>), dim3(nblocks), dim3(NT), 0, st,
                           a);
    return hipGetLastError();
}

// delayed-update windows; the measurement-only pass is the one-move kernel's
template <typename T, int N, int LPR>
static hipError_t launch_sweep_wdu_t(const SweepArgs &a, int nblocks, hipStream_t st)
{
    constexpr int NT = ((N * LPR + 63) / 64) * 64;
    constexpr size_t lds = wdu_lds_bytes<T, N, LPR>();
    if (a.measure_only) return launch_sweep_t<T, N, LPR, true>(a, nblocks, st);
    hipLaunchKernelGGL((sweep_kernel<T, N, LPR, true, false, true>), dim3(nblocks), dim3(NT), lds, st, a);
    return hipGetLastError();
}

template <typename OUT, int N, int LPR>
static hipError_t launch_gb_t(const GBArgs &a, int nwalkers, hipStream_t st)
{
    constexpr int NT = ((N * LPR + 63) / 64) * 64;
    hipLaunchKernelGGL((gbuild_kernel<OUT, N, LPR>), dim3(2 * nwalkers), dim3(NT), 0, st, a);
    return hipGetLastError();
}

template <typename OUT, int N>
static hipError_t launch_gb_mfma(const GBArgs &a, int nwalkers, hipStream_t st)
{
    hipLaunchKernelGGL((gbuild_mfma_kernel<OUT, N>), dim3(2 * nwalkers), dim3(N * 4), 0, st, a);
    return hipGetLastError();
}

struct SweepEntry {
    int prec, N, lpr, res;
    hipError_t (*fn)(const SweepArgs &, int, hipStream_t);
};
struct SweepWduEntry {
    int prec, N, lpr;
    hipError_t (*fn)(const SweepArgs &, int, hipStream_t);
};
struct GBEntry {
    int prec, N, lpr;
    hipError_t (*fn)(const GBArgs &, int, hipStream_t);
};

// first entry of a (prec, N) pair is the default launch shape
static const SweepEntry kSweep[] = {
    {0, 8, 8, 1, launch_sweep_t<double, 8, 8, true>},
    {0, 8, 4, 1, launch_sweep_t<double, 8, 4, true>},
    {0, 18, 2, 1, launch_sweep_t<double, 18, 2, true>},
    {0, 32, 4, 1, launch_sweep_t<double, 32, 4, true>},
    {0, 32, 2, 1, launch_sweep_t<double, 32, 2, true>},
    {0, 32, 8, 1, launch_sweep_t<double, 32, 8, true>},
    {0, 50, 2, 0, launch_sweep_t<double, 50, 2, false>},
    {0, 72, 4, 0, launch_sweep_t<double, 72, 4, false>},
    {0, 128, 2, 0, launch_sweep_t<double, 128, 2, false>},
    {0, 200, 2, 0, launch_sweep_t<double, 200, 2, false>},
    {1, 8, 8, 1, launch_sweep_t<float, 8, 8, true>},
    {1, 18, 2, 1, launch_sweep_t<float, 18, 2, true>},
    {1, 32, 4, 1, launch_sweep_t<float, 32, 4, true>},
    {1, 32, 2, 1, launch_sweep_t<float, 32, 2, true>},
    {1, 50, 2, 1, launch_sweep_t<float, 50, 2, true>},
    {1, 72, 4, 1, launch_sweep_t<float, 72, 4, true>},
    {1, 128, 4, 1, launch_sweep_t<float, 128, 4, true>},
    {1, 200, 2, 0, launch_sweep_t<float, 200, 2, false>},
};
static const SweepWduEntry kSweepWdu[] = {
    {1, 128, 4, launch_sweep_wdu_t<float, 128, 4>},
};
static const GBEntry kGB[] = {
    {0, 8, 8, launch_gb_t<double, 8, 8>},
    {0, 18, 2, launch_gb_t<double, 18, 2>},
    {0, 32, 4, launch_gb_mfma<double, 32>},
    {0, 50, 2, launch_gb_t<double, 50, 2>},
    {0, 72, 4, launch_gb_t<double, 72, 4>},
    {0, 128, 8, launch_gb_mfma<double, 128>},
    {0, 200, 2, launch_gb_t<double, 200, 2>},
    {1, 8, 8, launch_gb_t<float, 8, 8>},
    {1, 18, 2, launch_gb_t<float, 18, 2>},
    {1, 32, 4, launch_gb_mfma<float, 32>},
    {1, 50, 2, launch_gb_t<float, 50, 2>},
    {1, 72, 4, launch_gb_t<float, 72, 4>},
    {1, 128, 8, launch_gb_mfma<float, 128>},
    {1, 200, 2, launch_gb_t<float, 200, 2>},
};

}  // namespace gv

extern "C" hipError_t gv_launch_ss_reduce(const double *ss, const int32_t *idx, const int32_t *off,
                                          int G, size_t per, double *out, hipStream_t st)
{
    const dim3 grid((unsigned)((per + 255) / 256), (unsigned)G);
    hipLaunchKernelGGL(gv::ss_reduce_kernel, grid, dim3(256), 0, st, ss, idx, off, per, out);
    return hipGetLastError();
}

extern "C" int gv_sweep_supported(int prec, int N, int lpr, int *res, int *nthreads)
{
    for (const auto &e : gv::kSweep)
        if (e.prec == prec && e.N == N && (lpr == 0 || e.lpr == lpr)) {
            if (res) *res = e.res;
            if (nthreads) *nthreads = ((e.N * e.lpr + 63) / 64) * 64;
            return e.lpr;
        }
    return 0;
}

extern "C" int gv_sweep_shapes(int prec, int N, int *lpr, int max)
{
    int n = 0;
    for (const auto &e : gv::kSweep)
        if (e.prec == prec && e.N == N) {
            if (lpr && n < max) lpr[n] = e.lpr;
            ++n;
        }
    return n;
}

extern "C" hipError_t gv_launch_sweep(int prec, int N, int lpr, const gv::SweepArgs *a, int nblocks,
                                      hipStream_t st)
{
    for (const auto &e : gv::kSweep)
        if (e.prec == prec && e.N == N && e.lpr == lpr) return e.fn(*a, nblocks, st);
    return hipErrorInvalidValue;
}

extern "C" int gv_sweep_wdu_supported(int prec, int N, int lpr)
{
    for (const auto &e : gv::kSweepWdu)
        if (e.prec == prec && e.N == N && e.lpr == lpr) return 1;
    return 0;
}

extern "C" hipError_t gv_launch_sweep_wdu(int prec, int N, int lpr, const gv::SweepArgs *a, int nblocks,
                                          hipStream_t st)
{
    for (const auto &e : gv::kSweepWdu)
        if (e.prec == prec && e.N == N && e.lpr == lpr) return e.fn(*a, nblocks, st);
    return hipErrorInvalidValue;
}

extern "C" int gv_refresh_supported(int prec, int N)
{
    if (!gv_detinv_supported(N)) return 0;
    for (const auto &e : gv::kGB)
        if (e.prec == prec && e.N == N) return 1;
    return 0;
}

extern "C" hipError_t gv_launch_gbuild(int prec, int N, const gv::GBArgs *a, int nwalkers, hipStream_t st)
{
    for (const auto &e : gv::kGB)
        if (e.prec == prec && e.N == N) return e.fn(*a, nwalkers, st);
    return hipErrorInvalidValue;
}

namespace gv {

// One thread per walker; the rejection draw and the label pass touch only the walker's own
// site row (Ns ints) and part row, so the walkers are independent.  A draw that has not
// found n_up sites after 2^20 uniforms (probability far below 2^-1000 for n_up < Ns) is
// reported through *incomplete.
__global__ void __launch_bounds__(64) raninit_kernel(RIArgs a)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.W) return;
    if (a.status && a.status[2 * w] == 0 && a.status[2 * w + 1] == 0) return;
    int32_t *site = a.site + (size_t)w * a.Ns;
    int32_t *part = a.part + (size_t)w * 2 * a.nstride;
    for (int v = 0; v < a.Ns; ++v) site[v] = -1;
    uint32_t i = 0;
    int np = 0;
    while (np != a.n_up && i < (1u << 20)) {
        uint32_t c[4] = {i++, (uint32_t)a.attempt, (uint32_t)w, 2u};
        philox4x32_10_dev(c, a.key0, a.key1);
        const double u = ((double)c[0] + 0.5) * 0x1p-32;
        const int v = (int)(u * a.Ns);
        if (site[v] < 0) {
            site[v] = 0;
            ++np;
        }
    }
    if (np != a.n_up) *a.incomplete = 1;
    int cnt[2] = {0, 0};
    for (int v = 0; v < a.Ns; ++v) {
        const int fl = site[v] == 0 ? 0 : 1;
        const int lab = cnt[fl]++;
        site[v] = (fl << 16) | lab;
        if (lab < a.nstride) part[fl * a.nstride + lab] = v;
    }
    for (int fl = 0; fl < 2; ++fl)
        for (int lab = cnt[fl]; lab < a.nstride; ++lab) part[fl * a.nstride + lab] = 0;
}

}  // namespace gv

extern "C" hipError_t gv_launch_raninit(const gv::RIArgs *a, hipStream_t st)
{
    if (a->W <= 0 || a->Ns <= 0 || a->n_up < 0 || a->n_up >= a->Ns || a->n_up > a->nstride ||
        a->Ns - a->n_up > a->nstride || !a->site || !a->part || !a->incomplete)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(gv::raninit_kernel, dim3((a->W + 63) / 64), dim3(64), 0, st, *a);
    return hipGetLastError();
}
