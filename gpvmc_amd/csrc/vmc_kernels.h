// Kernel argument blocks shared by vmc_kernels.hip and the C-ABI layer (gpvmc_api.cpp).
#ifndef GPVMC_VMC_KERNELS_H
#define GPVMC_VMC_KERNELS_H
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gv {

enum : uint32_t { MEAS_PROJHEIS = 1u, MEAS_STAGMAGNZ = 2u, MEAS_OVERLAP = 4u, MEAS_STATSPIN = 8u };
constexpr int ACC_STRIDE = 12;

struct SweepArgs {
    void *G;                  // T2 [W][2][N][N], ratio tables G[fl][label][other label]
    int32_t *site;            // [W][Ns] (flavour << 16) | label
    int32_t *part;            // [W][2][N] site of each label
    double *logabs;           // [W][2] ln|det| per flavour
    double *phase;            // [W][2] arg det per flavour
    uint64_t *counter;        // [W] Philox step counter
    double *acc;              // [W][ACC_STRIDE]
    const void *occ;          // T2 [G][2][Ns][N]
    const int32_t *group;     // [W]
    const int32_t *order;     // [blocks] block -> walker (may be null)
    const int32_t *nbr;       // [Ns][4]
    const int4 *edge;         // [2Ns] (a, slot of b in nbr[a], b, slot of a in nbr[b])
    const double *uniforms;   // [W][n_steps][4] replay stream or null (Philox)
    int8_t *tr_noop;          // [W][n_steps] or null
    int8_t *tr_accept;
    double *tr_ratio;
    double *tr_eloc;          // [W][n_meas_slots] complex or null
    double *tr_mz;
    double *tr_logw;
    int64_t n_steps;
    int64_t meas;
    int64_t n_meas_slots;
    // register-resident kernel: this launch's first step and trace slot within the call
    // (walk segments between measurements), and the call's steps per walker
    int64_t st0, n_total, meas_slot;
    uint32_t mask;
    uint32_t key0, key1;
    int32_t L;
    int32_t measure_only;     // 1: no step, one measurement into the traces only
    double *ss;               // StatSpinStruct sums [W][SS_BLOCK(Ns)] or null
};

// StatSpinStruct sums of one walker (all FP64): szz[Ns][Ns] | smp[Ns][Ns] complex |
// spm diagonal [Ns] | measurement count.  spm off the diagonal is the transpose of smp
// (both receive -ratio of the same swap), spp and smm vanish for two flavours.
__host__ __device__ constexpr size_t ss_block(size_t Ns) { return 3 * Ns * Ns + Ns + 1; }

struct GBArgs {
    void *G;                  // OUT2 [W][2][N][N]
    const double2 *Ai;        // [W][2][N][N] FP64
    const int32_t *part;      // [W][2][N]
    const double2 *occ;       // [G][2][Ns][N] FP64
    const int32_t *group;     // [W]
    double *drift;            // [W][2] or null
};

// Channel (exciton) walk: runtime sizes n_up, n_dn, N_k; one workgroup per walker.
struct ExcArgs {
    void *Ai;                 // T2 [W][2][nmax][nmax] inverse of the current state, [label][slot]
    int32_t *site;            // [W][Ns] (flavour << 16) | label
    int32_t *part;            // [W][2][nmax] site of each label
    int32_t *cur;             // [W][2][nmax] orbital in each slot of the current state
    int32_t *cstate;          // [W] current channel state
    double *weight;           // [W] sum_k |a_k / a_c|^2 at the current configuration
    double *logabs;           // [W][2] ln|det| of the current state's matrices
    double *phase;            // [W][2]
    uint64_t *counter;        // [W]
    double *acc;              // [W][ACC_STRIDE]
    double *ho;               // [W][2 nk][nk] complex ProjHeis H and O sums
    const void *tab;          // T2 [G][2][Ns][Ns] orbital table Phi[fl][orbital][site]
    const void *tab64;        // double2, same layout
    const int32_t *group;     // [W]
    const int32_t *order;     // [blocks] block -> walker (may be null)
    const int32_t *nbr;       // [Ns][4]
    const int4 *edge;         // [2Ns]
    const int32_t *chan;      // [W] channel of each walker (may be null: channel 0)
    const int32_t *hop_n;     // [nch][nk][nk][2] orbital hops from state c to k per flavour (<= 2)
    const int32_t *hop_orb;   // [nch][nk][nk][2][2][2] (from, to) orbitals
    const int32_t *hop_new;   // [nch][nk][nk][2][2] index of the `to` orbital in new_orb of state c
    const int32_t *new_off;   // [nch][nk*2 + 1] offsets of new_orb per (c, fl)
    const int32_t *new_orb;   // orbitals some hop from state c fills, per (channel, c, fl)
    const int32_t *hopsign;   // [nch][nk][nk] State::HopSign of the path c -> k on state c
    const double *uniforms;
    int8_t *tr_noop;
    int8_t *tr_accept;
    double *tr_ratio;
    double *tr_logw;
    double *amp_out;          // [W][nk][2] ln|a_k|, arg a_k (mode 4)
    // register-resident variant: per walker the rows [Ai_fl; Q_fl] of both flavours,
    // Q_fl[q][f] = sum_r Phi_fl(qorb_fl[q], site r) Ai_fl[r][f] (the new orbitals' rows)
    void *M;                  // T2 [W][(rb0 + rb1) * 16][4 * cb] (flavour 0 row blocks first)
    int32_t *qorb;            // [W][2][max(nq0, nq1)] orbital of each Q row
    void *Q;                  // streamed kernel: T2 [W][2][max(nq0, nq1)][nmax] Q rows, [row][slot]
    void *Y;                  // streamed kernel measurement: T2 [W][2][nmax + nqm][nmax] of
                              // [Ai; Q] x Phi(current orbitals, sites of the other flavour's particles)
    const void *tabT;         // T2 [G][2][Ns sites][Ns orbitals] transposed orbital table
    void *zscr;               // T2 [W][16 * ceil(nmax / 16) targets][(rb0 + rb1) * 16] measurement tiles
    // register-resident kernel: measurements not yet folded into ho, per walker hb_cap
    // records of 2 nk + 1 values (a_k, then h_k, then the weight S in .x); a measurement
    // launch writes record hb_slot, gv_launch_excr_fold adds records 0 .. cnt-1 to ho
    double2 *hbuf;
    int32_t hb_cap, hb_slot;
    int32_t nq0, nq1;         // Q rows per flavour
    int32_t rb0, rb1, cb;     // row blocks of 16 per flavour, column blocks of 4
    int64_t n_steps;
    int64_t meas;
    int64_t n_meas_slots;
    // register-resident kernel: this launch's first step and trace slot within the call
    // (walk segments between measurements), and the call's steps per walker
    int64_t st0, n_total, meas_slot;
    uint32_t key0, key1;
    int32_t L, n_up, n_dn, nk;
    int32_t mode;             // 0 walk, 1 measurement only, 2 initial state choice, 3 weight only, 4 amplitudes,
                              // 5 measurement into the sums (register-resident kernel)
};

// Philox4x32-10 (Salmon et al., SC'11): counter (c0..c3), key (k0, k1).
__device__ __forceinline__ void philox4x32_10_dev(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c[0];
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        uint32_t lo1 = 0xCD9E8D57u * c[2];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// FP64 inversion with log-determinant (detinv_kernels.hip), one block per matrix.
// Modes 0/1: block 2w + fl inverts flavour fl of walker w, A(f, r) = Phi(orbital of slot f,
// site of label r) from the occupied-orbital table (mode 0, ground state: slot f holds
// orbital column f of occ) or from the full table and the slots' orbitals (mode 1,
// channels); out Ai[w][fl][label][slot] with row length nstride.  Mode 2: block b inverts
// the caller's column-major matrix A[b] (element (i, j) at i + j n) into the same layout.
struct DIArgs {
    int32_t mode;
    int32_t n_up, n_dn;       // matrix size of flavour 0 / 1 (mode 2: n_up for every matrix)
    int32_t nstride;          // modes 0/1: row length of part, cur and the outputs
    int32_t Ns;               // sites (modes 0/1)
    int32_t ld;               // leading dimension of `work` (streamed kernel, n > 128)
    const int32_t *part;      // [W][2][nstride] site of each label
    const int32_t *cur;       // [W][2][nstride] orbital in each slot (mode 1)
    const int32_t *group;     // [W]
    const double2 *occ;       // mode 0: [G][2][Ns][n]
    const double2 *tab;       // mode 1: [G][2][Ns][Ns] Phi[fl][orbital][site]
    const double2 *A;         // mode 2: [batch][n][n] column-major
    double2 *Ai64;            // out, FP64
    float2 *Ai32;             // modes 0/1: the same in FP32, or null
    double *logabs, *phase;   // [blocks] ln|det|, arg det
    int32_t *status;          // [blocks] 1 = singular (|det|^2 < 1e-9 or a zero pivot)
    double2 *work;            // [blocks][n][ld] working matrices of the streamed kernel, or null
};

// LatticeState::RanInit (LatticeState.cpp:37-82) on the device, one thread per walker:
// flavour 0 takes site floor(u Ns) of each uniform u until n_up distinct sites are taken,
// flavour 1 the rest; uniform i of walker w at attempt t is Philox4x32-10 (key seed, counter
// (i, t, w, 2)), u = (x0 + 1/2) 2^-32.  Labels in increasing site order (State::InitFock):
// site[w][v] = (flavour << 16) | label, part[w][fl * nstride + label] = v, unused part
// entries 0.  With status non-null only the walkers whose status pair is not (0, 0) draw.
// A walker whose draw stops at the cap of 2^20 uniforms sets *incomplete to 1.
struct RIArgs {
    int32_t W, Ns, nstride, n_up, n_dn, attempt;
    uint32_t key0, key1;
    const int32_t *status;    // [W][2] or null
    int32_t *site, *part;
    int32_t *incomplete;      // one int, cleared by the caller
};

}  // namespace gv

extern "C" {
hipError_t gv_launch_raninit(const gv::RIArgs *a, hipStream_t st);
int gv_sweep_supported(int prec, int N, int lpr, int *res, int *nthreads);
/* the lanes-per-row values of every sweep shape of (prec, N), default first; returns the count */
int gv_sweep_shapes(int prec, int N, int *lpr, int max);
/* delayed-update windows (several accepted moves per table update) for this shape */
int gv_sweep_wdu_supported(int prec, int N, int lpr);
hipError_t gv_launch_sweep_wdu(int prec, int N, int lpr, const gv::SweepArgs *a, int nblocks, hipStream_t st);
hipError_t gv_launch_sweep(int prec, int N, int lpr, const gv::SweepArgs *a, int nblocks,
                           hipStream_t st);
int gv_refresh_supported(int prec, int N);
hipError_t gv_launch_gbuild(int prec, int N, const gv::GBArgs *a, int nwalkers, hipStream_t st);
// out[g][e] = sum over the walkers idx[off[g] .. off[g+1]) of ss[walker][e], e < per
hipError_t gv_launch_exc(int prec, const gv::ExcArgs *a, int nblocks, hipStream_t st);
size_t gv_exc_lds_bytes(int Ns, int nmax, int nk, int nqm);
// Q rows of the streamed channel kernel from the FP64 inverse after Gauss-Jordan
hipError_t gv_launch_exc_qbuild(int prec, const gv::ExcArgs *a, const double2 *Ai64, int nwalkers, hipStream_t st);
// records the message returned by gpvmc_last_error and returns code (gpvmc_api.cpp)
int gv_fail(int code, const char *msg);
// FP64 inversion with log-determinant: 1 if n is supported, the work bytes per block
// (0 when the register-resident kernel covers n) and the work buffer's leading dimension
int gv_detinv_supported(int n);
size_t gv_detinv_work_bytes(int n);
int gv_detinv_ld(int n);
hipError_t gv_launch_detinv(const gv::DIArgs *a, int nblocks, hipStream_t st);
// register-resident channel kernel: the smallest registered shape (column blocks >= cb_min,
// row-block slots >= row_blocks) of a precision (1 = found)
int gv_excr_shape(int prec, int cb_min, int row_blocks, int *cb, int *rbw);
size_t gv_excr_lds_bytes(int prec, int L, int n_up, int n_dn, int nq0, int nq1, int rb0, int rb1, int cb, int nk);
size_t gv_excr_meas_lds_bytes(int prec, int L, int n_up, int n_dn, int nq0, int nq1, int rb0, int rb1, int cb, int nk);
hipError_t gv_launch_excr(int prec, int rbw, const gv::ExcArgs *a, int nblocks, hipStream_t st);
hipError_t gv_launch_excr_mbuild(int prec, const gv::ExcArgs *a, const double2 *Ai64, int nwalkers, hipStream_t st);
hipError_t gv_launch_excr_fold(const gv::ExcArgs *a, int cnt, int nwalkers, hipStream_t st);
hipError_t gv_launch_ss_reduce(const double *ss, const int32_t *idx, const int32_t *off, int G,
                               size_t per, double *out, hipStream_t st);
}

#endif
