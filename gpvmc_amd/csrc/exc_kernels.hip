// gfx950 kernels of the channel (exciton) walk: one workgroup per walker, runtime sizes.
//
// A channel is a set of N_k Slater states k; the walker's weight is w = sum_k |a_k|^2
// (LatticeStepper::trystep, LatticeStepper.cpp:81-95).  Each walker keeps the inverse
// Ai_fl of its "current" state c (the dominant state after the last accepted move, as
// LatticeStepper::step hops the wavefunction to argmax_k |a_k|^2, :102-113) in the slot
// arrangement the reference keeps (WaveFunction::Hop, State::Hop: an orbital replacement
// keeps its slot).  Rows of Ai are lattice particles (labels), columns are slots.
//
// Ratios (SlaterDeterminant::VirtUpdate, SlaterDeterminant.cpp:96-268, and
// linalg::DetUpdate, linalg.cpp:107-400).  A move of particle r to site s together with
// the replacement of the orbitals in slots f_m by o'_m (the hop path from c to k) changes
// one column and R rows of A; by the matrix determinant lemma the amplitude ratio is the
// determinant of the (1+R) x (1+R) capacitance matrix B with, for c = Phi(orbitals, s),
// y = Ai c, rho_m(r') = Phi(o'_m, site r'), and
// D_m = Phi(o'_m, s) - Phi(o'_m, site r) - Phi(o_m, s) + Phi(o_m, site r):
//   B00 = y[r],  B0n = Ai[r][f_n],
//   Bm0 = rho_m . y + D_m y[r] - Phi(o'_m, s),  Bmn = (rho_m Ai)[f_n] + D_m Ai[r][f_n].
// The crossing terms use the current A, so the ratios are exact (the reference uses the
// A of its last Init there, linalg.cpp:196,320,457).  a_k(x') / a_c(x) is the product of
// both flavours' determinants times HopSign(c -> k).  Without a lattice move B is the
// R x R block (rho_m Ai)[f_n].
//
// Accepted move to state k*: Woodbury, Ai <- Ai - U B^-1 V with U = [y - e_r, Ai[:, f_m]],
// V = [Ai[r, :]; (rho_m Ai) + D_m Ai[r, :] - e_{f_m}] (InvUpdate, linalg.cpp:402-573).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vmc_kernels.h"
#include "exc_common.h"

namespace gx {

using gv::ExcArgs;

using namespace gz;

// Explicit address spaces for the walk kernel's pointers: global memory (GAS) and LDS (LAS).
// Pointers kept in the walker's state are otherwise generic, every access through them a FLAT
// instruction, which also counts against the LDS counter and so makes each LDS access wait
// for the global loads in flight.
#define GAS __attribute__((address_space(1)))
#define LAS __attribute__((address_space(3)))
template <typename X> __device__ __forceinline__ X GAS *gptr(X *p) { return (X GAS *)p; }
// loads and stores of the two- and four-component vector types through global pointers go
// through clang's native vector types (the HIP vector types copy only from generic memory)
template <typename X> struct natv;
template <> struct natv<float2> { typedef float t __attribute__((ext_vector_type(2))); };
template <> struct natv<double2> { typedef double t __attribute__((ext_vector_type(2))); };
template <typename X> __device__ __forceinline__ X ldg(const X GAS *p)
{
    const typename natv<X>::t v = *reinterpret_cast<const typename natv<X>::t GAS *>(p);
    X r;
    r.x = v.x;
    r.y = v.y;
    return r;
}
template <typename X> __device__ __forceinline__ void stg(X GAS *p, X v)
{
    typename natv<X>::t n;
    n.x = v.x;
    n.y = v.y;
    *reinterpret_cast<typename natv<X>::t GAS *>(p) = n;
}
__device__ __forceinline__ int4 ldg4(const int4 GAS *p)
{
    typedef int i4 __attribute__((ext_vector_type(4)));
    const i4 v = *reinterpret_cast<const i4 GAS *>(p);
    int4 r;
    r.x = v.x;
    r.y = v.y;
    r.z = v.z;
    r.w = v.w;
    return r;
}

// LDS layout of one walker (dynamic shared memory; offsets computed by lds_carve)
struct Lds {
    int LAS *site;      // [Ns]   (flavour << 16) | label
    int LAS *part;      // [2][nmax] site of each label
    int LAS *cur;       // [2][nmax] orbital in each slot of the current state
    int LAS *slot;      // [2][Ns] slot of each orbital or -1
    int LAS *qrow;      // [2][Ns] Q row of each orbital or -1
    int LAS *qorb;      // [2][nqm] orbital of each Q row or -1
    int LAS *rep;       // [nqm] apply: slot whose unit row replaces a Q row, or -1
    int LAS *kqi;       // [nqm] apply: index of each Q row's orbital among the current state's new orbitals
    double LAS *vc;     // [2][nmax] complex: c = Phi(cur orbitals, target site)
    double LAS *vy;     // [2][nmax] complex: y = Ai c
    double LAS *kg;     // [2][nqm] complex: rho_m . y per new orbital of the current state
    double LAS *ra;     // [nk][2][2][2] complex: (rho_m Ai)[f_n] blocks of every state
    double LAS *alpha;  // [nk] complex: a_k(x') / a_c(x)
    double LAS *a0;     // [nk] complex: amplitudes at the current configuration
    double LAS *hk;     // [nk] complex: ProjHeis h_k
    double LAS *kap;    // apply: [nmax + nqm][3] complex update coefficients of one flavour's rows
    double LAS *dlt;    // apply: [nqm] complex delta of each Q row
    double LAS *V;      // apply: [3][nmax] complex V rows, then W = B^-1 V
    double LAS *red;    // [16]
    int LAS *ired;      // [16]
};

// complex elements of the apply coefficients' area, which the measurement reuses for its
// tiles of Phi and its bond list (two ints per bond, 2 Ns bonds)
__host__ __device__ inline size_t kap_elems(int Ns, int nmax, int nqm)
{
    size_t k = ((size_t)nmax + nqm) * 3;
    k = k > (size_t)Ns ? k : (size_t)Ns;
    // the measurement's 16-column tile of Phi ((nmax + 3) x 16 complex FP32) spans this area,
    // dlt and V
    const size_t t = 5 * (size_t)nmax + 24;
    return k > t ? k : t;
}

__host__ __device__ inline size_t exc_lds_bytes(int Ns, int nmax, int nk, int nqm)
{
    size_t b = 0;
    b += ((size_t)Ns + 2 * (size_t)nmax * 2 + 2 * (size_t)Ns * 2 + 2 * (size_t)nqm + 2 * (size_t)nqm) * 4;
    b = (b + 15) & ~(size_t)15;
    b += 16 * ((size_t)2 * nmax * 2 + (size_t)2 * nqm + (size_t)nk * 8 + nk + nk + nk +
               kap_elems(Ns, nmax, nqm) + nqm + (size_t)3 * nmax);
    b += 16 * 8 + 16 * 4;
    return b;
}

__device__ inline Lds lds_carve(char LAS *base, int Ns, int nmax, int nk, int nqm)
{
    Lds s;
    int LAS *ip = reinterpret_cast<int LAS *>(base);
    s.site = ip; ip += Ns;
    s.part = ip; ip += 2 * nmax;
    s.cur = ip; ip += 2 * nmax;
    s.slot = ip; ip += 2 * Ns;
    s.qrow = ip; ip += 2 * Ns;
    s.qorb = ip; ip += 2 * nqm;
    s.rep = ip; ip += nqm;
    s.kqi = ip; ip += nqm;
    size_t off = (reinterpret_cast<char LAS *>(ip) - base + 15) & ~(size_t)15;
    double LAS *dp = reinterpret_cast<double LAS *>(base + off);
    s.vc = dp; dp += 2 * (2 * nmax);
    s.vy = dp; dp += 2 * (2 * nmax);
    s.kg = dp; dp += 2 * (2 * nqm);
    s.ra = dp; dp += 2 * (nk * 8);
    s.alpha = dp; dp += 2 * nk;
    s.a0 = dp; dp += 2 * nk;
    s.hk = dp; dp += 2 * nk;
    s.kap = dp; dp += 2 * kap_elems(Ns, nmax, nqm);
    s.dlt = dp; dp += 2 * nqm;
    s.V = dp; dp += 2 * (3 * nmax);
    s.red = dp; dp += 16;
    s.ired = reinterpret_cast<int LAS *>(dp);
    return s;
}

// ln|product| of a mantissa/exponent pair per flavour (both flavours summed)
__device__ __forceinline__ double amp_log(const double LAS *m)
{
    return log(hypot(m[0], m[1])) + log(hypot(m[2], m[3])) + (m[4] + m[5]) * 0.69314718055994531;
}

template <typename T> struct Walker {
    using C = typename cpx<T>::t;
    const ExcArgs &a;
    Lds s;
    int w, tid, nt, Ns, n[2], nmax, nk, nq[2], nqm;
    size_t cb;  // channel of the walker times nk: base of its graph arrays
    C GAS *Ai[2];
    C GAS *Qr[2];   // Q rows: Q[q][f] = (rho_{qorb[q]} Ai)[f], kept up to date with Ai
    const C GAS *tab[2];
    const C GAS *tabT[2];  // [site][orbital]
    // thread 0: product of the accepted moves' determinant ratios per flavour since the
    // launch began, mantissa [0..3] and binary exponent [4..5] (LDS)
    double LAS *amp;

    __device__ __forceinline__ Walker(const ExcArgs &args, char *base, int wk)
        : a(args), w(wk), tid(threadIdx.x), nt(blockDim.x)
    {
        Ns = a.L * a.L;
        n[0] = a.n_up;
        n[1] = a.n_dn;
        nmax = n[0] > n[1] ? n[0] : n[1];
        nk = a.nk;
        nq[0] = a.nq0;
        nq[1] = a.nq1;
        nqm = nq[0] > nq[1] ? nq[0] : nq[1];
        if (nqm < 1) nqm = 1;
        cb = (size_t)(a.chan ? gptr(a.chan)[w] : 0) * nk;
        s = lds_carve((char LAS *)base, Ns, nmax, nk, nqm);
        Ai[0] = gptr(reinterpret_cast<C *>(a.Ai)) + (size_t)w * 2 * nmax * nmax;
        Ai[1] = Ai[0] + (size_t)nmax * nmax;
        Qr[0] = gptr(reinterpret_cast<C *>(a.Q)) + (size_t)w * 2 * nqm * nmax;
        Qr[1] = Qr[0] + (size_t)nqm * nmax;
        tab[0] = gptr(reinterpret_cast<const C *>(a.tab)) + (size_t)gptr(a.group)[w] * 2 * Ns * Ns;
        tab[1] = tab[0] + (size_t)Ns * Ns;
        tabT[0] = gptr(reinterpret_cast<const C *>(a.tabT)) + (size_t)gptr(a.group)[w] * 2 * Ns * Ns;
        tabT[1] = tabT[0] + (size_t)Ns * Ns;
    }
    // per-flavour members by a select rather than a runtime index, so that the walker's
    // state stays in registers
    __device__ int nfl(int fl) const { return fl ? n[1] : n[0]; }
    __device__ int nql(int fl) const { return fl ? nq[1] : nq[0]; }
    __device__ C GAS *aip(int fl) const { return fl ? Ai[1] : Ai[0]; }
    __device__ C GAS *qrp(int fl) const { return fl ? Qr[1] : Qr[0]; }
    __device__ const C GAS *tbp(int fl) const { return fl ? tab[1] : tab[0]; }
    __device__ const C GAS *tTp(int fl) const { return fl ? tabT[1] : tabT[0]; }
    __device__ zc phi(int fl, int orb, int site) const { return toz(ldg(tbp(fl) + (size_t)orb * Ns + site)); }
    __device__ zc ai(int fl, int r, int f) const { return toz(ldg(aip(fl) + (size_t)r * nmax + f)); }
    // Q row q (an orbital without a row reads as zero; the rows are the new orbitals of the
    // current state, which every lookup asks for)
    __device__ zc qv(int fl, int q, int f) const
    {
        return q >= 0 ? toz(ldg(qrp(fl) + (size_t)q * nmax + f)) : zmk(0, 0);
    }
    __device__ zc lz(const double LAS *p, int i) const { return zmk(p[2 * i], p[2 * i + 1]); }
    __device__ void sz(double LAS *p, int i, zc v) const { p[2 * i] = v.x; p[2 * i + 1] = v.y; }

    __device__ int cstate() const { return gptr(a.cstate)[w]; }
    __device__ size_t ck(int c, int k) const { return (cb + c) * nk + k; }
    __device__ int hop_n(int c, int k, int fl) const { return gptr(a.hop_n)[ck(c, k) * 2 + fl]; }
    __device__ int hop_from(int c, int k, int fl, int h) const { return gptr(a.hop_orb)[((ck(c, k) * 2 + fl) * 2 + h) * 2]; }
    __device__ int hop_to(int c, int k, int fl, int h) const { return gptr(a.hop_orb)[((ck(c, k) * 2 + fl) * 2 + h) * 2 + 1]; }
    __device__ int hop_new(int c, int k, int fl, int h) const  // index into the new orbitals of (c, fl)
    {
        return gptr(a.hop_new)[(ck(c, k) * 2 + fl) * 2 + h];
    }
    __device__ int hopsign(int c, int k) const { return gptr(a.hopsign)[ck(c, k)]; }
    // hop orbitals clamped to a valid index, for loads issued before hop_n is known (their
    // results are masked where h >= hop_n)
    __device__ int hop_from_c(int c, int k, int fl, int h) const { return oclamp(hop_from(c, k, fl, h)); }
    __device__ int hop_to_c(int c, int k, int fl, int h) const { return oclamp(hop_to(c, k, fl, h)); }
    __device__ int oclamp(int o) const { return o < 0 ? 0 : (o < Ns ? o : Ns - 1); }
    // new_off holds 2 nk + 1 offsets per channel
    __device__ const int32_t GAS *noff(int c, int fl) const { return gptr(a.new_off) + (cb / nk) * (2 * nk + 1) + c * 2 + fl; }
    __device__ int n_new(int c, int fl) const { return noff(c, fl)[1] - noff(c, fl)[0]; }
    __device__ int new_orb(int c, int fl, int m) const { return gptr(a.new_orb)[noff(c, fl)[0] + m]; }

    // RA blocks of every state relative to the current state (after any change of Ai or
    // of the current state): (rho_{o'_m} Ai)[f_n] is entry f_n of the Q row of o'_m
    __device__ __forceinline__ void build_ra()
    {
        const int c = cstate();
        // every load issued unconditionally (clamped indices), the entries outside the hop
        // count masked afterwards
        // a thread per state: its hop data once, then its eight Q-row entries together
        for (int k = tid; k < nk; k += nt) {
            int R[2], q[2][2], f[2][2];
#pragma unroll
            for (int fl = 0; fl < 2; ++fl) {
                R[fl] = hop_n(c, k, fl);
#pragma unroll
                for (int m = 0; m < 2; ++m) {
                    q[fl][m] = s.qrow[fl * Ns + hop_to_c(c, k, fl, m)];
                    f[fl][m] = s.slot[fl * Ns + hop_from_c(c, k, fl, m)];
                }
            }
            C v[2][2][2];
#pragma unroll
            for (int fl = 0; fl < 2; ++fl)
#pragma unroll
                for (int m = 0; m < 2; ++m)
#pragma unroll
                    for (int nn = 0; nn < 2; ++nn)
                        v[fl][m][nn] = ldg(qrp(fl) + (size_t)(q[fl][m] >= 0 ? q[fl][m] : 0) * nmax +
                                           (f[fl][nn] >= 0 ? f[fl][nn] : 0));
#pragma unroll
            for (int fl = 0; fl < 2; ++fl)
#pragma unroll
                for (int m = 0; m < 2; ++m)
#pragma unroll
                    for (int nn = 0; nn < 2; ++nn)
                        sz(s.ra, ((k * 2 + fl) * 2 + m) * 2 + nn,
                           m < R[fl] && nn < R[fl] && q[fl][m] >= 0 ? toz(v[fl][m][nn]) : zmk(0, 0));
        }
        __syncthreads();
    }

    // Ratios of every state for the swap moving up particle ci to site sb and down
    // particle cj to site sa (mv = true), or at the current configuration (mv = false):
    // fills alpha[k] and dets[k][fl]; leaves c, y, kg in LDS for apply().
    __device__ __forceinline__ void eval(bool mv, int ci, int sa, int sb, int cj)
    {
        const int c = cstate();
        if (mv) {
            for (int fl = 0; fl < 2; ++fl) {
                const int tgt = fl ? sa : sb;
                for (int f = tid; f < nfl(fl); f += nt)
                    sz(s.vc, fl * nmax + f, phi(fl, s.cur[fl * nmax + f], tgt));
            }
            __syncthreads();
            // y = Ai c and KG = rho . y: sixteen lanes per row (f or r = l16 + 16 j, j < 16
            // as nmax <= 256, a row's sixteen consecutive elements in one cache line), the
            // partial sums reduced across the sixteen lanes.  Loads are unconditional (lanes
            // past the row's end repeat its last element, which their product masks out) and
            // the next row's loads are issued before this row's products, so that two rows
            // per group stay in flight.
            const int l16 = tid & 15, g16 = tid >> 4, ng16 = nt >> 4;
            {
                const int rows = n[0] + n[1];
                auto ld = [&](int e, C (&x)[16]) {
                    const int ec = e < rows ? e : rows - 1;
                    const int fl = ec < n[0] ? 0 : 1, r = fl ? ec - n[0] : ec, nf = nfl(fl);
                    const C GAS *row = aip(fl) + (size_t)r * nmax;
#pragma unroll
                    for (int j = 0; j < 16; ++j) x[j] = ldg(row + (l16 + 16 * j < nf ? l16 + 16 * j : nf - 1));
                };
                auto fin = [&](int e, const C (&x)[16]) {
                    const int ec = e < rows ? e : rows - 1;
                    const int fl = ec < n[0] ? 0 : 1, r = fl ? ec - n[0] : ec, nf = nfl(fl);
                    zc t2[2] = {zmk(0, 0), zmk(0, 0)};
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const int f = l16 + 16 * j < nf ? l16 + 16 * j : nf - 1;
                        const zc cv = l16 + 16 * j < nf ? lz(s.vc, fl * nmax + f) : zmk(0, 0);
                        t2[j & 1] = zadd(t2[j & 1], zmul(toz(x[j]), cv));
                    }
                    zc t = zadd(t2[0], t2[1]);
                    for (int o = 8; o > 0; o >>= 1) {
                        t.x += __shfl_xor(t.x, o, 16);
                        t.y += __shfl_xor(t.y, o, 16);
                    }
                    if (l16 == 0 && e < rows) sz(s.vy, fl * nmax + r, t);
                };
                C xa[16], xb[16];
                ld(g16, xa);
                for (int e = g16; e < rows; e += 2 * ng16) {
                    ld(e + ng16, xb);
                    fin(e, xa);
                    ld(e + 2 * ng16, xa);
                    fin(e + ng16, xb);
                }
            }
            __syncthreads();
            // KG_m = rho_{o_m} . y over the particles, a thread per new orbital o_m of the current
            // state, from the transposed orbital table: for particle r every lane reads the
            // row of its site, at its own orbital's column, in blocks of 16 particles with the
            // next block's loads in flight (particles past a flavour's count repeat its last
            // one, masked out of the sums)
            {
                const int nn0 = n_new(c, 0), nn1 = n_new(c, 1), rows = nn0 + nn1;
                for (int e0 = 0; e0 < rows; e0 += nt) {
                    const int e = e0 + tid, ec = e < rows ? e : rows - 1;
                    const int fl = ec < nn0 ? 0 : 1, m = fl ? ec - nn0 : ec, nf = nfl(fl);
                    const C GAS *col = tTp(fl) + new_orb(c, fl, m);
                    zc t2[2] = {zmk(0, 0), zmk(0, 0)};
                    auto ld = [&](int r0, C (&x)[16]) {
#pragma unroll
                        for (int j = 0; j < 16; ++j)
                            x[j] = ldg(col + (size_t)s.part[fl * nmax + (r0 + j < nf ? r0 + j : nf - 1)] * Ns);
                    };
                    auto fin = [&](int r0, const C (&x)[16]) {
#pragma unroll
                        for (int j = 0; j < 16; ++j) {
                            const int r = r0 + j < nf ? r0 + j : nf - 1;
                            const zc yv = r0 + j < nf ? lz(s.vy, fl * nmax + r) : zmk(0, 0);
                            t2[j & 1] = zadd(t2[j & 1], zmul(toz(x[j]), yv));
                        }
                    };
                    C xa[16], xb[16];
                    ld(0, xa);
                    for (int r0 = 0; r0 < nmax; r0 += 32) {
                        ld(r0 + 16, xb);
                        fin(r0, xa);
                        ld(r0 + 32, xa);
                        fin(r0 + 16, xb);
                    }
                    if (e < rows) sz(s.kg, fl * nqm + m, zadd(t2[0], t2[1]));
                }
            }
            __syncthreads();
        }
        for (int k = tid; k < nk; k += nt) {
            zc tot = zmk((double)hopsign(c, k), 0.0);
            for (int fl = 0; fl < 2; ++fl) {
                const int R = hop_n(c, k, fl);
                // the capacitance matrix padded to 3 x 3 with the identity (constant indices);
                // every load is issued before R is known (clamped indices) and masked after
                zc B[9];
                eye3(B);
                if (!mv) {
#pragma unroll
                    for (int m = 0; m < 2; ++m)
#pragma unroll
                        for (int q = 0; q < 2; ++q)
                            if (m < R && q < R) B[m * 3 + q] = lz(s.ra, ((k * 2 + fl) * 2 + m) * 2 + q);
                } else {
                    const int r = fl ? cj : ci;
                    const int sold = fl ? sb : sa, snew = fl ? sa : sb;
                    int o1[2], o2[2], hn[2];
                    zc aq[2], p2s[2], p2r[2], p1s[2], p1r[2], kgv[2];
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        o1[m] = hop_from_c(c, k, fl, m);
                        o2[m] = hop_to_c(c, k, fl, m);
                        const int h = hop_new(c, k, fl, m);  // o'_m's place among the new orbitals
                        hn[m] = h < 0 ? 0 : (h < nqm ? h : nqm - 1);
                    }
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        const int f = s.slot[fl * Ns + o1[m]];
                        aq[m] = ai(fl, r, f >= 0 ? f : 0);
                        p2s[m] = phi(fl, o2[m], snew);
                        p2r[m] = phi(fl, o2[m], sold);
                        p1s[m] = phi(fl, o1[m], snew);
                        p1r[m] = phi(fl, o1[m], sold);
                        kgv[m] = lz(s.kg, fl * nqm + hn[m]);
                    }
                    B[0] = lz(s.vy, fl * nmax + r);
#pragma unroll
                    for (int q = 0; q < 2; ++q)
                        if (q < R) B[1 + q] = aq[q];
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        if (m >= R) continue;
                        const zc Dm = zadd(zsub(zsub(p2s[m], p2r[m]), p1s[m]), p1r[m]);
                        B[(1 + m) * 3] = zsub(zadd(kgv[m], zmul(Dm, B[0])), p2s[m]);
#pragma unroll
                        for (int q = 0; q < 2; ++q)
                            if (q < R)
                                B[(1 + m) * 3 + 1 + q] = zadd(lz(s.ra, ((k * 2 + fl) * 2 + m) * 2 + q),
                                                              zmul(Dm, B[1 + q]));
                    }
                }
                tot = zmul(tot, det3(B));
            }
            sz(s.alpha, k, tot);
        }
        __syncthreads();
    }

    // sum_k |alpha_k|^2 and the first k of largest |alpha_k|^2 (std::max_element)
    __device__ __forceinline__ void reduce_alpha(double &sum, int &kmax)
    {
        double ls = 0.0, bv = -1.0;
        int bk = nk;
        for (int k = tid; k < nk; k += nt) {
            const double v = znorm(lz(s.alpha, k));
            ls += v;
            if (v > bv) { bv = v; bk = k; }
        }
        for (int o = 32; o > 0; o >>= 1) {
            ls += __shfl_xor(ls, o);
            const double ov = __shfl_xor(bv, o);
            const int ok = __shfl_xor(bk, o);
            if (ov > bv || (ov == bv && ok < bk)) { bv = ov; bk = ok; }
        }
        const int lane = tid & 63, wid = tid >> 6;
        __syncthreads();
        if (lane == 0) { s.red[wid] = ls; s.red[8 + wid] = bv; s.ired[wid] = bk; }
        __syncthreads();
        sum = 0.0;
        bv = -1.0;
        bk = nk;
        for (int q = 0; q < (nt + 63) / 64; ++q) {
            sum += s.red[q];
            if (s.red[8 + q] > bv || (s.red[8 + q] == bv && s.ired[q] < bk)) { bv = s.red[8 + q]; bk = s.ired[q]; }
        }
        kmax = bk;
        __syncthreads();
    }

    // Moves the walker to state kn after eval(mv, ...): Woodbury update of both inverses
    // and of the Q rows, slot and Q-row bookkeeping, lattice hop.
    //   Ai' = Ai - U W,  W = B^-1 V,  U = [y - e_r, Ai[:, f_m]],
    //   V = [Ai[r, :]; Q_{o'_m} + D_m Ai[r, :] - e_{f_m}]   (InvUpdate, linalg.cpp:402-573)
    //   Q'_q = rho'_o Ai' with rho'_o = rho_o + delta_o e_r (delta_o = Phi(o, s) - Phi(o, site r)):
    //   Q'_q = Q_q + delta_o V_0 - sum_a kappa_a W_a,  kappa_0 = KG_q - Phi(o, s) + delta_o y[r],
    //   kappa_m = Q_q[f_m] + delta_o Ai[r][f_m].  The row of a filled orbital o'_m is taken
    //   over by the emptied orbital o_m, whose row before the move is e_{f_m} (it was
    //   occupied): kappa_0 = eps y[r], kappa_m = delta_{m,ms} + eps Ai[r][f_m], delta = eps =
    //   Phi(o_m, s) - Phi(o_m, site r).  Without a lattice move the delta terms vanish.
    __device__ __forceinline__ void apply(bool mv, int ci, int sa, int sb, int cj, int kn)
    {
        const int c = cstate();
        __shared__ zc s_binv[9];
        __shared__ zc s_det[2];
        for (int fl = 0; fl < 2; ++fl) {
            const int R = hop_n(c, kn, fl);
            const int r = fl ? cj : ci;
            const int sold = fl ? sb : sa, snew = fl ? sa : sb;
            const int m0 = mv ? 1 : 0, D = R + m0;
            const int nf = nfl(fl), nqf = nql(fl);
            if (D == 0) {
                if (tid == 0) s_det[fl] = zmk(1.0, 0.0);
                continue;
            }
            int fmv[2] = {0, 0}, o1v[2] = {0, 0}, o2v[2] = {0, 0};
            zc Dmv[2] = {zmk(0, 0), zmk(0, 0)};
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                if (m >= R) continue;
                o1v[m] = hop_from(c, kn, fl, m);
                o2v[m] = hop_to(c, kn, fl, m);
                fmv[m] = s.slot[fl * Ns + o1v[m]];
                if (mv)
                    Dmv[m] = zadd(zsub(zsub(phi(fl, o2v[m], snew), phi(fl, o2v[m], sold)), phi(fl, o1v[m], snew)),
                                  phi(fl, o1v[m], sold));
            }
            // V rows from the old inverse and Q rows
            for (int f = tid; f < nf; f += nt) {
                if (mv) sz(s.V, f, ai(fl, r, f));
#pragma unroll
                for (int m = 0; m < 2; ++m) {
                    if (m >= R) continue;
                    zc v = qv(fl, s.qrow[fl * Ns + o2v[m]], f);
                    if (mv) v = zadd(v, zmul(Dmv[m], ai(fl, r, f)));
                    if (f == fmv[m]) v.x -= 1.0;
                    sz(s.V, (m0 + m) * nmax + f, v);
                }
            }
            // KG of a Q row is the c-product eval formed for that orbital (kg, indexed by the
            // orbital's place among the current state's new orbitals)
            for (int q = tid; q < nqf; q += nt) s.kqi[q] = -1;
            __syncthreads();
            if (mv)
                for (int m = tid; m < n_new(c, fl); m += nt) {
                    const int q = s.qrow[fl * Ns + new_orb(c, fl, m)];
                    if (q >= 0) s.kqi[q] = m;
                }
            __syncthreads();
            // update coefficients of the inverse's rows (e < nf) and of the Q rows
            const zc yr = mv ? lz(s.vy, fl * nmax + r) : zmk(0, 0);
            for (int e = tid; e < nf + nqf; e += nt) {
                zc k3[3] = {zmk(0, 0), zmk(0, 0), zmk(0, 0)};
                if (e < nf) {
                    if (mv) {
                        zc y = lz(s.vy, fl * nmax + e);
                        if (e == r) y.x -= 1.0;
                        k3[0] = y;
                    }
#pragma unroll
                    for (int m = 0; m < 2; ++m)
                        if (m < R) k3[m0 + m] = ai(fl, e, fmv[m]);
                } else {
                    const int q = e - nf, o = s.qorb[fl * nqm + q];
                    zc dd = zmk(0, 0);
                    int rp = -1;
                    if (o >= 0) {
                        int ms = -1;
#pragma unroll
                        for (int m = 0; m < 2; ++m)
                            if (m < R && o2v[m] == o) ms = m;
                        if (ms >= 0) {
                            rp = ms ? fmv[1] : fmv[0];
                            const int o1s = ms ? o1v[1] : o1v[0];
                            const zc eps = mv ? zsub(phi(fl, o1s, snew), phi(fl, o1s, sold)) : zmk(0, 0);
                            dd = eps;
                            if (mv) k3[0] = zmul(eps, yr);
#pragma unroll
                            for (int m = 0; m < 2; ++m) {
                                if (m >= R) continue;
                                zc kv = zmk(m == ms ? 1.0 : 0.0, 0.0);
                                if (mv) kv = zadd(kv, zmul(eps, ai(fl, r, fmv[m])));
                                k3[m0 + m] = kv;
                            }
                        } else {
                            if (mv) {
                                // KG_q = Q_q . c = rho_o . y
                                zc kg = zmk(0, 0);
                                if (s.kqi[q] >= 0)
                                    kg = lz(s.kg, fl * nqm + s.kqi[q]);
                                else
                                    for (int f = 0; f < nf; ++f) kg = zadd(kg, zmul(qv(fl, q, f), lz(s.vc, fl * nmax + f)));
                                const zc pn = phi(fl, o, snew), dq = zsub(pn, phi(fl, o, sold));
                                dd = dq;
                                k3[0] = zadd(zsub(kg, pn), zmul(dq, yr));
                            }
#pragma unroll
                            for (int m = 0; m < 2; ++m) {
                                if (m >= R) continue;
                                zc kv = qv(fl, q, fmv[m]);
                                if (mv) kv = zadd(kv, zmul(dd, ai(fl, r, fmv[m])));
                                k3[m0 + m] = kv;
                            }
                        }
                    }
                    sz(s.dlt, q, dd);
                    s.rep[q] = rp;
                }
#pragma unroll
                for (int q3 = 0; q3 < 3; ++q3) sz(s.kap, e * 3 + q3, k3[q3]);
            }
            // capacitance matrix of state kn, rebuilt from the same inputs as eval
            // (padded to 3 x 3 with the identity: its inverse's leading D x D block is B^-1)
            if (tid == 0) {
                zc B[9];
                eye3(B);
                if (!mv) {
#pragma unroll
                    for (int m = 0; m < 2; ++m)
#pragma unroll
                        for (int q = 0; q < 2; ++q)
                            if (m < R && q < R) B[m * 3 + q] = lz(s.ra, ((kn * 2 + fl) * 2 + m) * 2 + q);
                } else {
                    B[0] = yr;
#pragma unroll
                    for (int q = 0; q < 2; ++q)
                        if (q < R) B[1 + q] = ai(fl, r, fmv[q]);
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        if (m >= R) continue;
                        const zc p2s = phi(fl, o2v[m], snew);
                        B[(1 + m) * 3] = zsub(zadd(lz(s.kg, fl * nqm + hop_new(c, kn, fl, m)), zmul(Dmv[m], B[0])), p2s);
#pragma unroll
                        for (int q = 0; q < 2; ++q)
                            if (q < R)
                                B[(1 + m) * 3 + 1 + q] = zadd(lz(s.ra, ((kn * 2 + fl) * 2 + m) * 2 + q), zmul(Dmv[m], B[1 + q]));
                    }
                }
                s_det[fl] = det3(B);
                inv3(B, s_binv);
            }
            __syncthreads();
            // W = B^-1 V in place; V_0 = Ai[r, :] is kept in this flavour's c vector, which
            // the coefficients above were the last to read
            for (int f = tid; f < nf; f += nt) {
                zc vv[3] = {zmk(0, 0), zmk(0, 0), zmk(0, 0)};
#pragma unroll
                for (int b2 = 0; b2 < 3; ++b2)
                    if (b2 < D) vv[b2] = lz(s.V, b2 * nmax + f);
                if (mv) sz(s.vc, fl * nmax + f, vv[0]);
#pragma unroll
                for (int a2 = 0; a2 < 3; ++a2) {
                    if (a2 >= D) continue;
                    zc wv = zmk(0, 0);
#pragma unroll
                    for (int b2 = 0; b2 < 3; ++b2)
                        if (b2 < D) wv = zadd(wv, zmul(s_binv[a2 * 3 + b2], vv[b2]));
                    sz(s.V, a2 * nmax + f, wv);
                }
            }
            __syncthreads();
            // rows of the inverse, then the live Q rows: a wave per row, lane l holding columns
            // f = min(l + 64 j, nf - 1) (nmax <= 256; lanes past the last column repeat it and
            // store the same value), W and V_0 kept in registers across the wave's rows,
            // arithmetic in the storage precision.  The loop body has no branch around a load
            // or a store, so that the next batch's loads stay in flight during this batch's
            // arithmetic and stores: a row without data (a Q row without orbital, or past the
            // last row) reads and writes the walker's measurement buffer, which every
            // measurement rewrites before reading it.
            {
                const int lane = tid & 63, nw = nt >> 6, rows = nf + nqf;
                int fc[4];
                C wr[3][4], v0[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int f = lane + 64 * j < nf ? lane + 64 * j : nf - 1;
                    fc[j] = f;
#pragma unroll
                    for (int a2 = 0; a2 < 3; ++a2) {
                        wr[a2][j].x = a2 < D ? (T)s.V[2 * (a2 * nmax + f)] : (T)0;
                        wr[a2][j].y = a2 < D ? (T)s.V[2 * (a2 * nmax + f) + 1] : (T)0;
                    }
                    v0[j].x = mv ? (T)s.vc[2 * (fl * nmax + f)] : (T)0;
                    v0[j].y = mv ? (T)s.vc[2 * (fl * nmax + f) + 1] : (T)0;
                }
                C GAS *dump = gptr(reinterpret_cast<C *>(a.Y)) + ((size_t)w * 2 + fl) * (nmax + nqm) * nmax;
                // batches of RB rows per wave (row = r0 + h nw), software-pipelined: the next
                // batch's loads are issued before this batch's arithmetic and stores
                constexpr int RB = 4;
                const int wid = __builtin_amdgcn_readfirstlane(tid >> 6), step = RB * nw;
                auto row_ptr = [&](int row) -> C GAS * {
                    if (row < nf) return aip(fl) + (size_t)row * nmax;
                    if (row < rows && s.qorb[fl * nqm + row - nf] >= 0) return qrp(fl) + (size_t)(row - nf) * nmax;
                    return dump + (size_t)(row % (nmax + nqm)) * nmax;
                };
                auto issue = [&](int r0, C GAS *(&p)[RB], C (&v)[RB][4]) {
#pragma unroll
                    for (int h = 0; h < RB; ++h) {
                        p[h] = row_ptr(r0 + h * nw);
#pragma unroll
                        for (int j = 0; j < 4; ++j) v[h][j] = ldg(p[h] + fc[j]);
                    }
                };
                auto finish = [&](int r0, C GAS *(&p)[RB], C (&v)[RB][4]) {
#pragma unroll
                    for (int h = 0; h < RB; ++h) {
                        const int row = r0 + h * nw;
                        const bool qr = row >= nf;
                        const int qi = qr ? (row - nf < nqm ? row - nf : nqm - 1) : 0;
                        const int kr = row < nf + nqm ? row : nf + nqm - 1;
                        C kp[3];
#pragma unroll
                        for (int a2 = 0; a2 < 3; ++a2) {
                            kp[a2].x = (T)s.kap[2 * (kr * 3 + a2)];
                            kp[a2].y = (T)s.kap[2 * (kr * 3 + a2) + 1];
                        }
                        const int rpl = qr ? s.rep[qi] : -1;
                        C dl;
                        dl.x = qr && mv ? (T)s.dlt[2 * qi] : (T)0;
                        dl.y = qr && mv ? (T)s.dlt[2 * qi + 1] : (T)0;
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            T ax = (T)0, ay = (T)0;
#pragma unroll
                            for (int a2 = 0; a2 < 3; ++a2) {
                                ax += kp[a2].x * wr[a2][j].x - kp[a2].y * wr[a2][j].y;
                                ay += kp[a2].x * wr[a2][j].y + kp[a2].y * wr[a2][j].x;
                            }
                            C nv = v[h][j];
                            if (rpl >= 0) {
                                nv.x = fc[j] == rpl ? (T)1 : (T)0;
                                nv.y = (T)0;
                            }
                            nv.x += dl.x * v0[j].x - dl.y * v0[j].y;
                            nv.y += dl.x * v0[j].y + dl.y * v0[j].x;
                            nv.x -= ax;
                            nv.y -= ay;
                            stg(p[h] + fc[j], nv);
                        }
                    }
                };
                C GAS *pa[RB], *pb[RB];
                C va[RB][4], vb[RB][4];
                issue(wid, pa, va);
                for (int r0 = wid; r0 < rows; r0 += 2 * step) {
                    issue(r0 + step, pb, vb);
                    finish(r0, pa, va);
                    issue(r0 + 2 * step, pa, va);
                    finish(r0 + step, pb, vb);
                }
            }
            __syncthreads();
        }
        // amplitude of the new current state (per flavour) and the slot / Q-row bookkeeping
        if (tid == 0) {
            for (int fl = 0; fl < 2; ++fl) {
                const zc d = s_det[fl];
                double mr = amp[2 * fl], mi = amp[2 * fl + 1], ex = amp[4 + fl];
                zamp_mul(mr, mi, ex, d.x, d.y);
                amp[2 * fl] = mr;
                amp[2 * fl + 1] = mi;
                amp[4 + fl] = ex;
            }
            for (int fl = 0; fl < 2; ++fl) {
                const int R = hop_n(c, kn, fl);
                for (int m = 0; m < R; ++m) {
                    const int o1 = hop_from(c, kn, fl, m), o2 = hop_to(c, kn, fl, m);
                    const int fm = s.slot[fl * Ns + o1];
                    s.slot[fl * Ns + o1] = -1;
                    s.slot[fl * Ns + o2] = fm;
                    s.cur[fl * nmax + fm] = o2;
                    const int q = s.qrow[fl * Ns + o2];
                    s.qrow[fl * Ns + o2] = -1;
                    s.qrow[fl * Ns + o1] = q;
                    if (q >= 0) s.qorb[fl * nqm + q] = o1;
                }
            }
            if (mv) {
                s.site[sb] = ci;
                s.site[sa] = (1 << 16) | cj;
                s.part[ci] = sb;
                s.part[nmax + cj] = sa;
            }
            gptr(a.cstate)[w] = kn;
        }
        __syncthreads();
    }
};

// mode 0: walk n_steps; 1: measurement only (traces, no accumulators); 2: initial state
// choice (the state of largest |a_k|^2 at the configuration, vmc.cpp:243-254) and weight
// bookkeeping after an init or a refresh of the inverse; 3: weight bookkeeping only;
// 4: amplitudes of every state into amp_out
template <typename T>
__global__ void __launch_bounds__(256) exc_kernel(ExcArgs a)
{
    extern __shared__ __align__(16) char smem[];
    const int w = a.order ? gptr(a.order)[blockIdx.x] : (int)blockIdx.x;
    Walker<T> wk(a, smem, w);
    __shared__ double s_amp[6];
    wk.amp = (double LAS *)s_amp;
    if (threadIdx.x == 0) {
        wk.amp[0] = 1.0; wk.amp[1] = 0.0; wk.amp[2] = 1.0; wk.amp[3] = 0.0;
        wk.amp[4] = 0.0; wk.amp[5] = 0.0;
    }
    const int tid = wk.tid, nt = wk.nt, Ns = wk.Ns, nmax = wk.nmax, nk = wk.nk;
    auto &s = wk.s;
    const int nqm = wk.nqm;
    for (int i = tid; i < Ns; i += nt) {
        s.site[i] = gptr(a.site)[(size_t)w * Ns + i];
        s.slot[i] = -1;
        s.slot[Ns + i] = -1;
        s.qrow[i] = -1;
        s.qrow[Ns + i] = -1;
    }
    for (int i = tid; i < 2 * nmax; i += nt) {
        s.part[i] = gptr(a.part)[(size_t)w * 2 * nmax + i];
        s.cur[i] = gptr(a.cur)[(size_t)w * 2 * nmax + i];
    }
    for (int i = tid; i < 2 * nqm; i += nt) s.qorb[i] = gptr(a.qorb)[(size_t)w * 2 * nqm + i];
    __syncthreads();
    for (int i = tid; i < 2 * nmax; i += nt) {
        const int fl = i / nmax, f = i % nmax;
        if (f < wk.nfl(fl)) s.slot[fl * Ns + s.cur[i]] = f;
    }
    for (int i = tid; i < 2 * nqm; i += nt) {
        const int fl = i / nqm, q = i % nqm;
        if (q < wk.nql(fl) && s.qorb[i] >= 0) s.qrow[fl * Ns + s.qorb[i]] = q;
    }
    __syncthreads();
    wk.build_ra();
    double S;  // sum_k |a_k / a_c|^2 at the current configuration
    int kmax;
    wk.eval(false, 0, 0, 0, 0);
    wk.reduce_alpha(S, kmax);
    if (a.mode == 2) {
        if (kmax != wk.cstate() && kmax < nk) {
            wk.apply(false, 0, 0, 0, 0, kmax);
            wk.build_ra();
            wk.eval(false, 0, 0, 0, 0);
            wk.reduce_alpha(S, kmax);
        }
    }
    int64_t n_noop = 0, n_acc = 0, n_rej = 0;
    double sacc[6] = {0, 0, 0, 0, 0, 0};  // [0] n_meas, [1] sum ln w
    const uint64_t ctr0 = gptr(a.counter)[w];
    int64_t mslot = 0;
    auto measure = [&](bool acc_on) {
        // ProjHeis::measure on the channel (ProjHeis.cpp:40-154): h_k = sum over bonds of
        // +a_k/4 (parallel) or -a_k/4 - alpha_k(swap)/2 (antiparallel).  The swap ratios of
        // every bond come from Y_fl = [Ai_fl; Q_fl] x Phi_fl(current orbitals, sites of the
        // other flavour's particles): y[r] of moving particle r onto the site of particle t
        // is Y[r][t], and rho_m . y is Y[n_fl + row of o'_m][t].
        wk.eval(false, 0, 0, 0, 0);
        double S0;
        int kk;
        wk.reduce_alpha(S0, kk);
        for (int k = tid; k < nk; k += nt) wk.sz(s.a0, k, wk.lz(s.alpha, k));
        using CT = typename cpx<T>::t;
        const int rows_max = nmax + nqm;
        // the chunk of Phi columns, then the bond list, live where apply keeps its coefficients
        double LAS *cbuf = s.kap;
        const int CH = (6 * nmax + 4 * nqm) / nmax < 8 ? (6 * nmax + 4 * nqm) / nmax : 8;
        if constexpr (sizeof(T) == 4) {
            // Q rows of Y on MFMA (v_mfma_f32_16x16x4f32): per 16 targets a tile of
            // Phi(current orbitals, target sites) in LDS (the B operand, [f][target], zero past
            // the last orbital and target), each wave a 16-row block of Q at a time (the A
            // operand straight from HBM, lane l holding row l % 16 and column l / 16 of each
            // 16 x 4 step, eight steps' loads in flight); the complex product as four real
            // products, Yr = Qr Phir - Qi Phii, Yi = Qr Phii + Qi Phir.  Output element v of
            // lane l is row 4 (l / 16) + v, target l % 16.
            typedef float v4 __attribute__((ext_vector_type(4)));
            const int lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nt >> 6;
            const int li = lane & 15, kq = lane >> 4;
            float LAS *ct = reinterpret_cast<float LAS *>(cbuf);
            for (int fl = 0; fl < 2; ++fl) {
                const int nf = wk.nfl(fl), no = wk.nfl(1 - fl), nqf = wk.nql(fl), nf4 = (nf + 3) & ~3;
                CT GAS *Yf = gptr(reinterpret_cast<CT *>(a.Y)) + ((size_t)w * 2 + fl) * rows_max * nmax;
                const int nqb = (nqf + 15) >> 4, nks = nf4 >> 2;
                for (int t0 = 0; t0 < no; t0 += 16) {
                    __syncthreads();
                    for (int e = tid; e < nf4 * 16; e += nt) {
                        const int f = e >> 4, tg = t0 + (e & 15);
                        CT v;
                        v.x = v.y = 0.f;
                        if (f < nf && tg < no)
                            v = ldg(wk.tTp(fl) + (size_t)s.part[(1 - fl) * nmax + tg] * Ns + s.cur[fl * nmax + f]);
                        ct[2 * e] = v.x;
                        ct[2 * e + 1] = v.y;
                    }
                    __syncthreads();
                    for (int qb = wid; qb < nqb; qb += nw) {
                        const int q = qb * 16 + li;
                        const CT GAS *arow = wk.qrp(fl) + (size_t)(q < nqf ? q : nqf - 1) * nmax;
                        v4 zr = {0.f, 0.f, 0.f, 0.f}, zi = {0.f, 0.f, 0.f, 0.f};
                        auto lda = [&](int k0, CT (&x)[8]) {
#pragma unroll
                            for (int j = 0; j < 8; ++j) {
                                const int f = 4 * (k0 + j) + kq;
                                x[j] = ldg(arow + (f < nf ? f : nf - 1));
                            }
                        };
                        auto mm = [&](int k0, const CT (&x)[8]) {
#pragma unroll
                            for (int j = 0; j < 8; ++j) {
                                const int f = 4 * (k0 + j) + kq;
                                const int fb = f < nf4 ? f : nf4 - 1;
                                const float bx = f < nf4 ? ct[2 * (fb * 16 + li)] : 0.f;
                                const float by = f < nf4 ? ct[2 * (fb * 16 + li) + 1] : 0.f;
                                zr = __builtin_amdgcn_mfma_f32_16x16x4f32(x[j].x, bx, zr, 0, 0, 0);
                                zr = __builtin_amdgcn_mfma_f32_16x16x4f32(x[j].y, -by, zr, 0, 0, 0);
                                zi = __builtin_amdgcn_mfma_f32_16x16x4f32(x[j].x, by, zi, 0, 0, 0);
                                zi = __builtin_amdgcn_mfma_f32_16x16x4f32(x[j].y, bx, zi, 0, 0, 0);
                            }
                        };
                        CT xa[8], xb[8];
                        lda(0, xa);
                        for (int k0 = 0; k0 < nks; k0 += 16) {
                            lda(k0 + 8, xb);
                            mm(k0, xa);
                            lda(k0 + 16, xa);
                            mm(k0 + 8, xb);
                        }
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int row = qb * 16 + 4 * kq + v, tg = t0 + li;
                            if (row < nqf && tg < no) {
                                CT y;
                                y.x = zr[v];
                                y.y = zi[v];
                                stg(Yf + (size_t)(nf + row) * nmax + tg, y);
                            }
                        }
                    }
                }
            }
        } else {
        // Q rows of Y: a thread per row, the row read in blocks of 16 elements with the next
        // block's loads in flight (unconditional loads, the last element repeated past the
        // row's end and masked out of the products)
        for (int fl = 0; fl < 2; ++fl) {
            const int nf = wk.nfl(fl), no = wk.nfl(1 - fl), nqf = wk.nql(fl);
            CT GAS *Yf = gptr(reinterpret_cast<CT *>(a.Y)) + ((size_t)w * 2 + fl) * rows_max * nmax;
            for (int t0 = 0; t0 < no; t0 += CH) {
                __syncthreads();
                for (int e = tid; e < nf * CH; e += nt) {
                    const int f = e / CH, j = e - f * CH, t = t0 + j;
                    wk.sz(cbuf, e, t < no ? wk.phi(fl, s.cur[fl * nmax + f], s.part[(1 - fl) * nmax + t]) : zmk(0, 0));
                }
                __syncthreads();
                for (int q = tid; q < nqf; q += nt) {
                    const CT GAS *rowp = wk.qrp(fl) + (size_t)q * nmax;
                    zc acc[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) acc[j] = zmk(0, 0);
                    auto ld = [&](int f0, CT (&x)[16]) {
#pragma unroll
                        for (int j = 0; j < 16; ++j) x[j] = ldg(rowp + (f0 + j < nf ? f0 + j : nf - 1));
                    };
                    auto mac = [&](int f0, const CT (&x)[16]) {
#pragma unroll
                        for (int j = 0; j < 16; ++j) {
                            const int f = f0 + j < nf ? f0 + j : nf - 1;
                            const zc m = f0 + j < nf ? toz(x[j]) : zmk(0, 0);
#pragma unroll
                            for (int jj = 0; jj < 8; ++jj)
                                acc[jj] = zadd(acc[jj], zmul(m, wk.lz(cbuf, f * CH + (jj < CH ? jj : CH - 1))));
                        }
                    };
                    CT xa[16], xb[16];
                    ld(0, xa);
                    for (int f0 = 0; f0 < nf; f0 += 32) {
                        ld(f0 + 16, xb);
                        mac(f0, xa);
                        ld(f0 + 32, xa);
                        mac(f0 + 16, xb);
                    }
                    for (int j = 0; j < 8; ++j)
                        if (j < CH && t0 + j < no) {
                            CT v;
                            v.x = (T)acc[j].x;
                            v.y = (T)acc[j].y;
                            stg(Yf + (size_t)(nf + q) * nmax + t0 + j, v);
                        }
                }
            }
        }
        }
        __syncthreads();
        // antiparallel bonds in bond order (a ballot prefix sum per round of nt bonds):
        // bl[2 b] = ci | cj << 16, bl[2 b + 1] = sa | sb << 16 (up particle ci at site sa,
        // down particle cj at site sb)
        __shared__ int s_wcnt[16];
        int LAS *bl = reinterpret_cast<int LAS *>(cbuf);
        int nb = 0;
        {
            const int lane = tid & 63, wid = tid >> 6, nwv = (nt + 63) >> 6;
            for (int e0 = 0; e0 < 2 * Ns; e0 += nt) {
                const int e = e0 + tid;
                bool ap = false;
                int v0 = 0, v1 = 0;
                if (e < 2 * Ns) {
                    const int4 eg = ldg4(gptr(a.edge) + e);
                    const int vi = eg.x, vj = eg.z;
                    const int ii = s.site[vi], ij = s.site[vj];
                    ap = (ii >> 16) != (ij >> 16);
                    if ((ii >> 16) == 0) { v0 = (ii & 0xFFFF) | ((ij & 0xFFFF) << 16); v1 = vi | (vj << 16); }
                    else { v0 = (ij & 0xFFFF) | ((ii & 0xFFFF) << 16); v1 = vj | (vi << 16); }
                }
                const unsigned long long bm = __ballot(ap);
                const int pre = __builtin_amdgcn_mbcnt_hi((unsigned)(bm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bm, 0u));
                if (lane == 0) s_wcnt[wid] = __popcll(bm);
                __syncthreads();
                int off = nb, tot = 0;
                for (int q = 0; q < nwv; ++q) {
                    if (q < wid) off += s_wcnt[q];
                    tot += s_wcnt[q];
                }
                if (ap) {
                    bl[2 * (off + pre)] = v0;
                    bl[2 * (off + pre) + 1] = v1;
                }
                nb += tot;
                __syncthreads();
            }
        }
        // the inverse rows' entries Y[r][t] the bonds read: y[r] of moving particle r (flavour
        // fl) onto the site of the other flavour's particle t, a sixteen-lane group per (bond,
        // flavour)
        {
            const int l16 = tid & 15, g16 = tid >> 4, ng16 = nt >> 4;
            for (int e = g16; e < 2 * nb; e += ng16) {
                const int b = e >> 1, fl = e & 1, nf = wk.nfl(fl);
                const int v0 = bl[2 * b], v1 = bl[2 * b + 1];
                const int r = fl ? v0 >> 16 : v0 & 0xFFFF, t = fl ? v0 & 0xFFFF : v0 >> 16;
                const int snew = fl ? v1 & 0xFFFF : v1 >> 16;
                const CT GAS *row = wk.aip(fl) + (size_t)r * nmax;
                CT x[16], pv[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int f = l16 + 16 * j < nf ? l16 + 16 * j : nf - 1;
                    x[j] = ldg(row + f);
                    pv[j] = ldg(wk.tbp(fl) + (size_t)s.cur[fl * nmax + f] * Ns + snew);
                }
                zc t2[2] = {zmk(0, 0), zmk(0, 0)};
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    if (l16 + 16 * j < nf) t2[j & 1] = zadd(t2[j & 1], zmul(toz(x[j]), toz(pv[j])));
                zc tt = zadd(t2[0], t2[1]);
                for (int o = 8; o > 0; o >>= 1) {
                    tt.x += __shfl_xor(tt.x, o, 16);
                    tt.y += __shfl_xor(tt.y, o, 16);
                }
                if (l16 == 0) {
                    CT v;
                    v.x = (T)tt.x;
                    v.y = (T)tt.y;
                    stg(gptr(reinterpret_cast<CT *>(a.Y)) + ((size_t)w * 2 + fl) * rows_max * nmax + (size_t)r * nmax + t, v);
                }
            }
        }
        __syncthreads();
        // h_k: a thread per state, the state's hop data held in registers across the bonds
        const int c = wk.cstate();
        const int npar = 2 * Ns - nb;
        for (int k = tid; k < nk; k += nt) {
            const zc a0 = wk.lz(s.a0, k);
            int Rv[2], o1v[2][2], o2v[2][2], fqv[2][2], qrv[2][2];
            zc rav[2][2][2];
#pragma unroll
            for (int fl = 0; fl < 2; ++fl) {
                Rv[fl] = wk.hop_n(c, k, fl);
#pragma unroll
                for (int m = 0; m < 2; ++m) {
                    const bool on = m < Rv[fl];
                    o1v[fl][m] = wk.hop_from_c(c, k, fl, m);
                    o2v[fl][m] = wk.hop_to_c(c, k, fl, m);
                    const int fq = s.slot[fl * Ns + o1v[fl][m]];
                    fqv[fl][m] = fq >= 0 ? fq : 0;
                    qrv[fl][m] = on ? s.qrow[fl * Ns + o2v[fl][m]] : -1;
#pragma unroll
                    for (int q = 0; q < 2; ++q) rav[fl][m][q] = wk.lz(s.ra, ((k * 2 + fl) * 2 + m) * 2 + q);
                }
            }
            const double hsg = (double)wk.hopsign(c, k);
            zc h = zmk(0, 0);
            // the bond's loads (all issued before the hop counts select, clamped indices), and
            // the next bond's issued before this one's determinants
            struct BondIn {
                CT y0[2], aq[2][2], p2s[2][2], p2r[2][2], p1s[2][2], p1r[2][2], kq[2][2];
            };
            auto bld = [&](int b, BondIn &in) {
                const int bc = b < nb ? b : nb - 1;
                const int v0 = bl[2 * bc], v1 = bl[2 * bc + 1];
                const int ci = v0 & 0xFFFF, cj = v0 >> 16, sa = v1 & 0xFFFF, sb = v1 >> 16;
#pragma unroll
                for (int fl = 0; fl < 2; ++fl) {
                    const int nf = wk.nfl(fl);
                    const int r = fl ? cj : ci, t = fl ? ci : cj;
                    const int sold = fl ? sb : sa, snew = fl ? sa : sb;
                    const CT GAS *Yf = gptr(reinterpret_cast<const CT *>(a.Y)) + ((size_t)w * 2 + fl) * rows_max * nmax;
                    const CT GAS *tf = wk.tbp(fl);
                    in.y0[fl] = ldg(Yf + (size_t)r * nmax + t);
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        in.aq[fl][m] = ldg(wk.aip(fl) + (size_t)r * nmax + fqv[fl][m]);
                        in.p2s[fl][m] = ldg(tf + (size_t)o2v[fl][m] * Ns + snew);
                        in.p2r[fl][m] = ldg(tf + (size_t)o2v[fl][m] * Ns + sold);
                        in.p1s[fl][m] = ldg(tf + (size_t)o1v[fl][m] * Ns + snew);
                        in.p1r[fl][m] = ldg(tf + (size_t)o1v[fl][m] * Ns + sold);
                        const int qr = qrv[fl][m];
                        in.kq[fl][m] = ldg(Yf + (size_t)(nf + (qr >= 0 ? qr : 0)) * nmax + t);
                    }
                }
            };
            auto bfin = [&](const BondIn &in) {
                zc tot = zmk(hsg, 0.0);
#pragma unroll
                for (int fl = 0; fl < 2; ++fl) {
                    const int R = Rv[fl];
                    zc B[9];  // padded to 3 x 3 with the identity
                    eye3(B);
                    B[0] = toz(in.y0[fl]);
#pragma unroll
                    for (int q = 0; q < 2; ++q)
                        if (q < R) B[1 + q] = toz(in.aq[fl][q]);
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        if (m >= R) continue;
                        const zc p2s = toz(in.p2s[fl][m]);
                        const zc Dm = zadd(zsub(zsub(p2s, toz(in.p2r[fl][m])), toz(in.p1s[fl][m])), toz(in.p1r[fl][m]));
                        const zc kq = qrv[fl][m] >= 0 ? toz(in.kq[fl][m]) : zmk(0, 0);
                        B[(1 + m) * 3] = zsub(zadd(kq, zmul(Dm, B[0])), p2s);
#pragma unroll
                        for (int q = 0; q < 2; ++q)
                            if (q < R) B[(1 + m) * 3 + 1 + q] = zadd(rav[fl][m][q], zmul(Dm, B[1 + q]));
                    }
                    tot = zmul(tot, det3(B));
                }
                h = zmk(h.x - 0.25 * a0.x - 0.5 * tot.x, h.y - 0.25 * a0.y - 0.5 * tot.y);
            };
            if (nb > 0) {
                BondIn ia, ib;
                bld(0, ia);
                for (int b = 0; b < nb; b += 2) {
                    bld(b + 1, ib);
                    bfin(ia);
                    bld(b + 2, ia);
                    if (b + 1 < nb) bfin(ib);
                }
            }
            // parallel bonds: +a_k / 4 each
            h = zmk(h.x + 0.25 * npar * a0.x, h.y + 0.25 * npar * a0.y);
            wk.sz(s.hk, k, h);
        }
        __syncthreads();
        const double lw = 2.0 * (gptr(a.logabs)[2 * w] + gptr(a.logabs)[2 * w + 1] + amp_log(wk.amp)) + log(S0);
        if (acc_on) {
            double2 GAS *HO = gptr(reinterpret_cast<double2 *>(a.ho)) + (size_t)w * 2 * nk * nk;
            // H and O sums in batches of four elements per thread, the next batch's loads
            // issued before this batch's stores (elements past the end repeat the last one,
            // which its own thread stores with the value it computed)
            const int ne = nk * nk;
            constexpr int HB = 4;
            auto hld = [&](int e0, double2 (&hs)[HB], double2 (&os)[HB]) {
#pragma unroll
                for (int j = 0; j < HB; ++j) {
                    const int e = e0 + j * nt < ne ? e0 + j * nt : ne - 1;
                    hs[j] = ldg(HO + e);
                    os[j] = ldg(HO + (size_t)nk * nk + e);
                }
            };
            auto hst = [&](int e0, double2 (&hs)[HB], double2 (&os)[HB]) {
#pragma unroll
                for (int j = 0; j < HB; ++j) {
                    const int e = e0 + j * nt;
                    if (e >= ne) continue;
                    const int k1 = e / nk, k2 = e - k1 * nk;
                    const zc a1 = wk.lz(s.a0, k1);
                    const zc h2 = wk.lz(s.hk, k2), a2 = wk.lz(s.a0, k2);
                    const zc hv = zconjmul(a1, h2), ov = zconjmul(a1, a2);  // conj(a_k1) h_k2, conj(a_k1) a_k2
                    hs[j].x += hv.x / S0;
                    hs[j].y += hv.y / S0;
                    os[j].x += ov.x / S0;
                    os[j].y += ov.y / S0;
                    stg(HO + e, hs[j]);
                    stg(HO + (size_t)nk * nk + e, os[j]);
                }
            };
            double2 ha[HB], oa[HB], hb[HB], ob[HB];
            const int stride = HB * nt;
            hld(tid, ha, oa);
            for (int e0 = tid; e0 < ne; e0 += 2 * stride) {
                hld(e0 + stride, hb, ob);
                hst(e0, ha, oa);
                hld(e0 + 2 * stride, ha, oa);
                hst(e0 + stride, hb, ob);
            }
            if (tid == 0) { sacc[0] += 1; sacc[1] += lw; }
        }
        if (tid == 0 && a.tr_logw) gptr(a.tr_logw)[(size_t)w * a.n_meas_slots + mslot] = lw;
        ++mslot;
        __syncthreads();
    };
    if (a.mode == 4) {
        // amplitudes of every state at the current configuration, as ln|a_k| and arg a_k
        const double la = gptr(a.logabs)[2 * w] + gptr(a.logabs)[2 * w + 1] + amp_log(wk.amp);
        const double ph = gptr(a.phase)[2 * w] + gptr(a.phase)[2 * w + 1] + atan2(wk.amp[1], wk.amp[0]) + atan2(wk.amp[3], wk.amp[2]);
        for (int k = tid; k < nk; k += nt) {
            const zc al = wk.lz(s.alpha, k);
            const double nrm = znorm(al);
            gptr(a.amp_out)[((size_t)w * nk + k) * 2] = nrm > 0.0 ? la + 0.5 * log(nrm) : -INFINITY;
            gptr(a.amp_out)[((size_t)w * nk + k) * 2 + 1] = remainder(ph + atan2(al.y, al.x), 2 * M_PI);
        }
    } else if (a.mode == 1) {
        measure(false);
    } else if (a.mode == 0) {
        for (int64_t st = 0; st < a.n_steps; ++st) {
            // proposal: 4 uniforms per step, as the ground-state kernel
            int v1, ed;
            double u3;
            if (a.uniforms) {
                const double GAS *up = gptr(a.uniforms) + ((size_t)w * a.n_steps + st) * 4;
                v1 = (int)(Ns * up[0]);
                v1 = v1 < Ns ? v1 : Ns - 1;
                ed = (int)(4 * up[1]);
                ed = ed < 4 ? ed : 3;
                u3 = up[3];
            } else {
                const uint64_t sc = ctr0 + (uint64_t)st;
                uint32_t cc[4] = {(uint32_t)sc, (uint32_t)(sc >> 32), (uint32_t)w, 0u};
                gv::philox4x32_10_dev(cc, a.key0, a.key1);
                v1 = (int)(((uint64_t)Ns * (2ull * cc[0] + 1ull)) >> 33);
                ed = (int)(cc[1] >> 30);
                u3 = ((double)cc[3] + 0.5) * 2.3283064365386963e-10;
            }
            const int v2 = gptr(a.nbr)[v1 * 4 + ed];
            const int i1 = s.site[v1], i2 = s.site[v2];
            const bool noop = (i1 >> 16) == (i2 >> 16);
            double ratio = 1.0;
            bool acc;
            if (noop) {
                acc = u3 < 1.0;
                ++n_noop;
            } else {
                int ci, cj, sa, sb;
                if ((i1 >> 16) == 0) { ci = i1 & 0xFFFF; sa = v1; cj = i2 & 0xFFFF; sb = v2; }
                else { ci = i2 & 0xFFFF; sa = v2; cj = i1 & 0xFFFF; sb = v1; }
                wk.eval(true, ci, sa, sb, cj);
                double S2;
                int kn;
                wk.reduce_alpha(S2, kn);
                ratio = S2 / S;
                acc = ratio > 1.0 ? (u3 < 1.0) : (u3 < ratio);  // MetroMC::YesNo
                if (acc) {
                    const zc ak = wk.lz(s.alpha, kn);
                    wk.apply(true, ci, sa, sb, cj, kn);
                    S = S2 / znorm(ak);
                    wk.build_ra();
                    ++n_acc;
                } else {
                    ++n_rej;
                }
            }
            if (tid == 0) {
                const size_t ti = (size_t)w * a.n_steps + st;
                if (a.tr_noop) gptr(a.tr_noop)[ti] = noop ? 1 : 0;
                if (a.tr_accept) gptr(a.tr_accept)[ti] = acc ? 1 : 0;
                if (a.tr_ratio) gptr(a.tr_ratio)[ti] = ratio;
            }
            if (a.meas > 0 && (a.meas == 1 || st % a.meas == 0)) measure(true);
        }
    }
    // write back
    __syncthreads();
    for (int i = tid; i < Ns; i += nt) gptr(a.site)[(size_t)w * Ns + i] = s.site[i];
    for (int i = tid; i < 2 * nmax; i += nt) {
        gptr(a.part)[(size_t)w * 2 * nmax + i] = s.part[i];
        gptr(a.cur)[(size_t)w * 2 * nmax + i] = s.cur[i];
    }
    for (int i = tid; i < 2 * nqm; i += nt) gptr(a.qorb)[(size_t)w * 2 * nqm + i] = s.qorb[i];
    if (tid == 0) {
        // ln|det| and arg det per flavour += the launch's product of ratios
        for (int fl = 0; fl < 2; ++fl) {
            const double mr = wk.amp[2 * fl], mi = wk.amp[2 * fl + 1];
            gptr(a.logabs)[2 * w + fl] += log(hypot(mr, mi)) + wk.amp[4 + fl] * 0.69314718055994531;
            gptr(a.phase)[2 * w + fl] = remainder(gptr(a.phase)[2 * w + fl] + atan2(mi, mr), 2 * M_PI);
        }
        gptr(a.weight)[w] = S;
        if (a.mode == 0) {
            gptr(a.counter)[w] = ctr0 + (uint64_t)a.n_steps;
            double GAS *ac = gptr(a.acc) + (size_t)w * gv::ACC_STRIDE;
            ac[3] += sacc[0];
            ac[7] += (double)a.n_steps;
            ac[8] += (double)n_acc;
            ac[9] += (double)n_rej;
            ac[10] += (double)n_noop;
            ac[11] += sacc[1];
        }
    }
}

}  // namespace gx

extern "C" hipError_t gv_launch_exc(int prec, const gv::ExcArgs *a, int nblocks, hipStream_t st)
{
    const int Ns = a->L * a->L;
    const int nmax = a->n_up > a->n_dn ? a->n_up : a->n_dn;
    const int nqm = a->nq0 > a->nq1 ? (a->nq0 > 0 ? a->nq0 : 1) : (a->nq1 > 0 ? a->nq1 : 1);
    if (!a->Q || !a->qorb || !a->Y || !a->tabT) return hipErrorInvalidValue;
    if (nmax > 256) return hipErrorInvalidValue;  // the update loop holds four column chunks of 64 per lane
    const size_t lds = gx::exc_lds_bytes(Ns, nmax, a->nk, nqm) + 256;
    if (lds > 160 * 1024 - 4096) return hipErrorInvalidValue;
    if (prec == 0)
        hipLaunchKernelGGL(gx::exc_kernel<double>, dim3(nblocks), dim3(256), lds, st, *a);
    else
        hipLaunchKernelGGL(gx::exc_kernel<float>, dim3(nblocks), dim3(256), lds, st, *a);
    return hipGetLastError();
}

extern "C" size_t gv_exc_lds_bytes(int Ns, int nmax, int nk, int nqm)
{
    return gx::exc_lds_bytes(Ns, nmax, nk, nqm < 1 ? 1 : nqm) + 256;
}

// Q rows of every walker's current state from the FP64 inverse: one block per (walker,
// flavour); qorb[q] = the q-th new orbital of the current state (-1 past them) and
// Q[q][f] = sum_r Phi(qorb[q], site r) Ai64[r][f] in the working precision, a complex
// (nq x n) x (n x n) product in FP64.  Output tiles of 64 x 64, each thread a 4 x 4
// micro-tile; the factors pass through LDS in chunks of 16 along r.
template <typename T>
__global__ void __launch_bounds__(256) exc_qbuild_kernel(gv::ExcArgs a, const double2 *Ai64)
{
    using C = typename gz::cpx<T>::t;
    constexpr int TQ = 64, TF = 64, KC = 16;
    __shared__ double2 s_rho[KC][TQ + 1];  // Phi(orbital of row q, site of label r), [r][q]
    __shared__ double2 s_ai[KC][TF];       // Ai64[r][f]
    __shared__ int s_orb[256], s_site[256];
    const int w = blockIdx.x >> 1, fl = blockIdx.x & 1;
    const int tid = threadIdx.x;
    const int Ns = a.L * a.L;
    const int nmax = a.n_up > a.n_dn ? a.n_up : a.n_dn;
    const int nqm = a.nq0 > a.nq1 ? (a.nq0 > 0 ? a.nq0 : 1) : (a.nq1 > 0 ? a.nq1 : 1);
    const int n = fl ? a.n_dn : a.n_up, nq = fl ? a.nq1 : a.nq0;
    const int c = a.cstate[w];
    const size_t chb = (size_t)(a.chan ? a.chan[w] : 0);
    const int32_t *off = a.new_off + chb * (2 * a.nk + 1) + c * 2 + fl;
    const int o0 = off[0], nnew = off[1] - off[0];
    const int nrow = nq < nnew ? nq : nnew;  // rows with an orbital; the rest are zero
    int32_t *qo = a.qorb + ((size_t)w * 2 + fl) * nqm;
    for (int q = tid; q < nqm; q += blockDim.x) qo[q] = (q < nq && q < nnew) ? a.new_orb[o0 + q] : -1;
    const int32_t *part = a.part + ((size_t)w * 2 + fl) * nmax;
    for (int i = tid; i < 256; i += blockDim.x) {
        s_orb[i] = i < nrow ? a.new_orb[o0 + i] : 0;
        s_site[i] = i < n ? part[i] : 0;
    }
    __syncthreads();
    const double2 *tab = reinterpret_cast<const double2 *>(a.tab64) + ((size_t)a.group[w] * 2 + fl) * Ns * Ns;
    const double2 *ai = Ai64 + ((size_t)w * 2 + fl) * nmax * nmax;
    C *Q = reinterpret_cast<C *>(a.Q) + ((size_t)w * 2 + fl) * nqm * nmax;
    const int tq = (tid / 16) * 4, tf = (tid % 16) * 4;  // micro-tile origin inside the output tile
    for (int q0 = 0; q0 < nq; q0 += TQ) {
        for (int f0 = 0; f0 < n; f0 += TF) {
            double2 acc[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = make_double2(0.0, 0.0);
            for (int k0 = 0; k0 < n; k0 += KC) {
                // stage: 64 x 16 of each factor, four elements per thread
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int idx = tid + 256 * e;  // 0 .. 1023
                    const int qq = idx / KC, rr = idx % KC;
                    const int q = q0 + qq, r = k0 + rr;
                    s_rho[rr][qq] = (q < nrow && r < n) ? tab[(size_t)s_orb[q] * Ns + s_site[r]] : make_double2(0.0, 0.0);
                    const int rk = idx / TF, ff = idx % TF;
                    const int r2 = k0 + rk, f = f0 + ff;
                    s_ai[rk][ff] = (r2 < n && f < n) ? ai[(size_t)r2 * nmax + f] : make_double2(0.0, 0.0);
                }
                __syncthreads();
#pragma unroll 4
                for (int kk = 0; kk < KC; ++kk) {
                    double2 x[4], y[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) x[i] = s_rho[kk][tq + i];
#pragma unroll
                    for (int j = 0; j < 4; ++j) y[j] = s_ai[kk][tf + j];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            acc[i][j].x = fma(x[i].x, y[j].x, acc[i][j].x);
                            acc[i][j].x = fma(-x[i].y, y[j].y, acc[i][j].x);
                            acc[i][j].y = fma(x[i].x, y[j].y, acc[i][j].y);
                            acc[i][j].y = fma(x[i].y, y[j].x, acc[i][j].y);
                        }
                }
                __syncthreads();
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int q = q0 + tq + i, f = f0 + tf + j;
                    if (q < nq && f < n) {
                        C v;
                        v.x = (T)acc[i][j].x;
                        v.y = (T)acc[i][j].y;
                        Q[(size_t)q * nmax + f] = v;
                    }
                }
        }
    }
}

extern "C" hipError_t gv_launch_exc_qbuild(int prec, const gv::ExcArgs *a, const double2 *Ai64, int nwalkers,
                                           hipStream_t st)
{
    if (!a->Q || !a->qorb) return hipErrorInvalidValue;
    if (a->nq0 > 256 || a->nq1 > 256 || a->n_up > 256 || a->n_dn > 256) return hipErrorInvalidValue;
    if (prec == 0)
        hipLaunchKernelGGL(exc_qbuild_kernel<double>, dim3(2 * nwalkers), dim3(256), 0, st, *a, Ai64);
    else
        hipLaunchKernelGGL(exc_qbuild_kernel<float>, dim3(2 * nwalkers), dim3(256), 0, st, *a, Ai64);
    return hipGetLastError();
}

