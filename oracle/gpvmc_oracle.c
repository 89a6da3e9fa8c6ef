/* gpvmc oracle: plain-C FP64 restatement of the reference's Metropolis VMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libgpvmc_hip.so, never part of the product.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Each function cites the reference file:line it restates.  How each one is pinned
 * (tests/test_oracle.py):
 *   or_neighbours/or_edges  reference-built harness dump (lattice mode), L=4,6,8
 *   or_mbz/or_orbitals      harness dump of WaveFunction::MatEl at L=4,8 and three
 *                           parameter sets; unitarity Phi Phi^H = Ns I (closed form)
 *   or_detinv               numpy.linalg det/inv on random complex matrices
 *   or_ratio/or_sm_update   brute-force det(A')/det(A) and inv(A') with numpy
 *   or_raninit/or_labels    harness RanInit configurations from the glibc rand stream
 *   or_init_walker          harness init amplitude and inverse (DetInv)
 *   or_walk                 harness per-step weight ratios, moves, snapshots and final
 *                           inverse over thousands of steps at L=4 and L=8
 *   or_eloc                 exact spin-basis H psi / psi at L=4 for every configuration
 *   or_stagmagnz            closed form on Neel and random configurations
 *   or_statspinstruct       exact <S^z S^z> and <S^+ S^-> at L=4 (exact enumeration);
 *                           sum rules
 *   or_exc_amps             numpy determinants of the same matrices; ratios a_k / a_0
 *                           equal to second-quantised site-basis amplitudes at L=4; the
 *                           reference's slot arrangement times its HopSign (fixtures)
 *   or_exc_walk             per-configuration h_k / a_k equal to the exact local
 *                           (H psi_k) / psi_k at L=4 (trans and long); the chain follows
 *                           the same proposal/accept code as or_walk
 */
#include "gpvmc_oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#define IDX(fl, a, b, n) ((size_t)(fl) * (size_t)(n) * (size_t)(n) + (size_t)(a) * (size_t)(n) + (size_t)(b))

static size_t umod(int x, int n)
{
    /* linalg::mod, linalg.h:165-168 */
    return x >= 0 ? (size_t)x % (size_t)n : (size_t)((n - abs(x) % n) % n);
}

/* SquareLattice ctor (SquareLattice.cpp:22-27) creates edges with x outer, y inner:
 * first (x,y)->(x,y+1), then (x,y)->(x+1,y); Edge ctor (Lattice.h:79-84) appends the edge
 * to both endpoints, so a vertex's edge list is in edge creation order. */
void or_edges(int L, int32_t *edges)
{
    int e = 0;
    for (int x = 0; x < L; ++x) {
        for (int y = 0; y < L; ++y) {
            edges[2 * e] = y * L + x;
            edges[2 * e + 1] = (int)umod(y + 1, L) * L + x;
            ++e;
            edges[2 * e] = y * L + x;
            edges[2 * e + 1] = y * L + (int)umod(x + 1, L);
            ++e;
        }
    }
}

void or_neighbours(int L, int32_t *nbr)
{
    int Ns = L * L;
    int *cnt = calloc((size_t)Ns, sizeof(int));
    int32_t *edges = malloc(sizeof(int32_t) * 4 * (size_t)Ns);
    or_edges(L, edges);
    for (int e = 0; e < 2 * Ns; ++e) {
        int a = edges[2 * e], b = edges[2 * e + 1];
        nbr[4 * a + cnt[a]++] = b;  /* Edge::GetOther(first) = second */
        nbr[4 * b + cnt[b]++] = a;
    }
    free(cnt);
    free(edges);
}

/* StagFluxWaveFunction::inmbz (StagFluxWaveFunction.cpp:89-101) */
static int inmbz(int L, double bcx, double bcy, int kx, int ky)
{
    double kkx = ((double)kx + bcx / 2.0) / L;
    double kky = ((double)ky + bcy / 2.0) / L;
    double eps = 1e-9;
    return (kkx >= 0 && kkx < 1 && kky >= 0 && kky < 1) &&
           (kkx + kky - 0.5 < 0 || kkx + kky - 1.5 >= -eps || kky - kkx - 0.5 > eps ||
            kky - kkx + 0.5 <= eps);
}

/* Orbital enumeration (StagFluxWaveFunction.cpp:24-41): kx outer, ky inner, two bands
 * per magnetic-zone momentum.  Returns the number of orbitals found (must be Ns). */
int or_mbz(int L, double bcx, double bcy, int32_t *fock2qn)
{
    int idx = 0;
    for (int kx = 0; kx < L; ++kx)
        for (int ky = 0; ky < L; ++ky)
            if (inmbz(L, bcx, bcy, kx, ky)) {
                fock2qn[3 * idx] = kx; fock2qn[3 * idx + 1] = ky; fock2qn[3 * idx + 2] = 0;
                ++idx;
                fock2qn[3 * idx] = kx; fock2qn[3 * idx + 1] = ky; fock2qn[3 * idx + 2] = 1;
                ++idx;
            }
    return idx;
}

/* delta, neelk, omega, Uk, Vk: StagFluxWaveFunction.cpp:121-163.  `sig` is +1 for
 * flavour 0 and -1 for flavour 1 (the flavour index is passed as `bool up`, :144). */
static cplx sf_delta(double phi, const double *k)
{
    return 0.5 * (cos(k[0]) * cexp(I * phi) + cos(k[1]) * cexp(-I * phi));
}

static double sf_neelk(double neel, double ne, const double *k)
{
    if (ne == 0.0) return neel;
    return neel * pow(0.5 * (pow(cos(k[0]), 2) + pow(cos(k[1]), 2)), ne);
}

static void sf_uv(double phi, double neel, double ne, const double *k, int sig, int band,
                  cplx *U, cplx *V)
{
    cplx d = sf_delta(phi, k);
    double nk = sf_neelk(neel, ne, k);
    double om = sqrt(nk * nk + creal(d) * creal(d) + cimag(d) * cimag(d));
    double ad = cabs(d);
    if (band == 0) {
        *U = csqrt(0.5 * (1 + sig * nk / om));
        *V = d / ad * csqrt(0.5 * (1 - sig * nk / om));
    } else {
        *U = -conj(d) / ad * csqrt(0.5 * (1 - sig * nk / om));
        *V = csqrt(0.5 * (1 + sig * nk / om));
    }
}

/* Full orbital table Phi_fl(f, r) (WaveFunction::init_matrices, WaveFunction.cpp:27-37;
 * StagFluxWaveFunction::matrix_element, StagFluxWaveFunction.cpp:51-79).
 * table[((fl*Ns + f)*Ns + r)*2 + {0,1}].  `phi` is in radians (vmc.cpp:120 multiplies the
 * command-line value by pi).  Returns 0, or -1 on a NaN orbital (the reference aborts). */
int or_orbitals(int L, double phi, double neel, double ne, double bcx, double bcy,
                double *table)
{
    int Ns = L * L;
    int32_t *qn = malloc(sizeof(int32_t) * 3 * (size_t)Ns);
    if (or_mbz(L, bcx, bcy, qn) != Ns) { free(qn); return -2; }
    cplx *t = (cplx *)table;
    int bad = 0;
    for (int fl = 0; fl < 2; ++fl) {
        int sig = fl == 0 ? 1 : -1;
        for (int f = 0; f < Ns; ++f) {
            double k[2];
            k[0] = (qn[3 * f] + bcx / 2.0) * 2 * M_PI / (double)L;
            k[1] = (qn[3 * f + 1] + bcy / 2.0) * 2 * M_PI / (double)L;
            int band = qn[3 * f + 2];
            cplx U, V;
            sf_uv(phi, neel, ne, k, sig, band, &U, &V);
            for (int r = 0; r < Ns; ++r) {
                double ry = (double)(r / L);
                double rx = (double)r - ry * L;
                int odd = ((int)(rx + ry)) % 2;
                cplx out = sqrt(2.0) * cexp(I * (k[0] * rx + k[1] * ry)) *
                           ((double)(1 - odd) * U + (double)odd * V);
                if (isnan(creal(out)) || isnan(cimag(out))) bad = 1;
                t[((size_t)fl * Ns + f) * Ns + r] = out;
            }
        }
    }
    free(qn);
    return bad ? -1 : 0;
}

/* Ground state occupies every band-0 orbital: wavefunction particle p sits in orbital 2p
 * (StagFluxGroundState.cpp:11-14).  occ[(fl*Ns + r)*N + p] = Phi_fl(2p, r). */
void or_ground_occ(int L, const double *table, double *occ)
{
    int Ns = L * L, N = Ns / 2;
    const cplx *t = (const cplx *)table;
    cplx *o = (cplx *)occ;
    for (int fl = 0; fl < 2; ++fl)
        for (int r = 0; r < Ns; ++r)
            for (int p = 0; p < N; ++p)
                o[((size_t)fl * Ns + r) * N + p] = t[((size_t)fl * Ns + 2 * p) * Ns + r];
}

/* linalg::DetInv (linalg.cpp:9-44): LU with partial pivoting (the zgetrf pivot is the
 * largest |re|+|im| in the column), det = sign(perm) * prod U_ii, then the inverse.  The
 * matrix is singular when |det|^2 < 1e-9 (:23); the inverse is then not formed.
 * A and Ai are N x N column-major in the reference's sense (A[r*N+f] = A(f, r)). */
int or_detinv(int N, const double *Ain, double *Aiout, double *logabs, double *phase)
{
    const cplx *A = (const cplx *)Ain;
    cplx *Ai = (cplx *)Aiout;
    cplx *lu = malloc(sizeof(cplx) * (size_t)N * N);
    int *piv = malloc(sizeof(int) * (size_t)N);
    memcpy(lu, A, sizeof(cplx) * (size_t)N * N);
    /* column-major: element (i, j) at i + j*N */
#define LU(i, j) lu[(size_t)(i) + (size_t)(j) * N]
    double la = 0.0, ph = 0.0;
    int sgn = 1, zero = 0;
    for (int j = 0; j < N; ++j) {
        int p = j;
        double best = -1.0;
        for (int i = j; i < N; ++i) {
            double v = fabs(creal(LU(i, j))) + fabs(cimag(LU(i, j)));
            if (v > best) { best = v; p = i; }
        }
        piv[j] = p;
        if (p != j) {
            sgn = -sgn;
            for (int c = 0; c < N; ++c) { cplx tmp = LU(j, c); LU(j, c) = LU(p, c); LU(p, c) = tmp; }
        }
        cplx d = LU(j, j);
        if (d == 0) { zero = 1; continue; }
        la += log(cabs(d));
        ph += carg(d);
        for (int i = j + 1; i < N; ++i) LU(i, j) /= d;
        for (int c = j + 1; c < N; ++c) {
            cplx f = LU(j, c);
            if (f != 0)
                for (int i = j + 1; i < N; ++i) LU(i, c) -= LU(i, j) * f;
        }
    }
    if (sgn < 0) ph += M_PI;
    ph = remainder(ph, 2 * M_PI);
    *logabs = zero ? -INFINITY : la;
    *phase = ph;
    int singular = zero || (2 * la < log(1e-9));
    if (!singular) {
        /* solve A X = I column by column */
        for (int c = 0; c < N; ++c) {
            cplx *x = &Ai[(size_t)c * N];
            for (int i = 0; i < N; ++i) x[i] = (i == c) ? 1.0 : 0.0;
            for (int j = 0; j < N; ++j)
                if (piv[j] != j) { cplx tmp = x[j]; x[j] = x[piv[j]]; x[piv[j]] = tmp; }
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < i; ++j) x[i] -= LU(i, j) * x[j];
            for (int i = N - 1; i >= 0; --i) {
                for (int j = i + 1; j < N; ++j) x[i] -= LU(i, j) * x[j];
                x[i] /= LU(i, i);
            }
        }
    }
#undef LU
    free(lu);
    free(piv);
    return singular;
}

/* Ratio of one column replacement, det(A')/det(A) = e_ci^T Ai C: the element of Ai*C that
 * linalg::DetUpdate keeps (linalg.cpp:143-154) with rc=1, rr=0 (a single lattice hop of
 * particle ci in one flavour). */
void or_ratio(int N, const double *Aiin, int ci, const double *colin, double *q)
{
    const cplx *Ai = (const cplx *)Aiin, *col = (const cplx *)colin;
    cplx s = 0;
    for (int p = 0; p < N; ++p) s += Ai[ci + (size_t)p * N] * col[p];
    q[0] = creal(s);
    q[1] = cimag(s);
}

/* Sherman-Morrison update of Ai after replacing column ci of A by `col`: linalg::InvUpdate
 * with rc=1, rr=0 (linalg.cpp:402-561): AiC = Ai*C; B = AiC[ci]; Bi = 1/B; AiC[ci] -= 1;
 * Ai -= AiC (Bi * row_ci(Ai)). */
void or_sm_update(int N, double *Aiin, int ci, const double *colin)
{
    cplx *Ai = (cplx *)Aiin;
    const cplx *col = (const cplx *)colin;
    cplx *aic = malloc(sizeof(cplx) * (size_t)N);
    cplx *mem = malloc(sizeof(cplx) * (size_t)N);
    for (int i = 0; i < N; ++i) {
        cplx s = 0;
        for (int p = 0; p < N; ++p) s += Ai[i + (size_t)p * N] * col[p];
        aic[i] = s;
    }
    cplx bi = 1.0 / aic[ci];
    aic[ci] -= 1.0;
    for (int p = 0; p < N; ++p) mem[p] = bi * Ai[ci + (size_t)p * N];
    for (int p = 0; p < N; ++p)
        for (int i = 0; i < N; ++i) Ai[i + (size_t)p * N] -= aic[i] * mem[p];
    free(aic);
    free(mem);
}

/* LatticeState::RanInit for two flavours at half filling with one internal state per
 * site (LatticeState.cpp:37-82): flavour 0 draws v = floor(u*Ns) until N distinct free
 * sites are taken; flavour 1 fills the remaining sites.  spins[v] = flavour at v.
 * Returns the number of uniforms consumed, or -1 if `u` ran out. */
int or_raninit(int L, const double *u, int64_t nu, int32_t *spins)
{
    int Ns = L * L, N = Ns / 2;
    int64_t used = 0;
    for (int v = 0; v < Ns; ++v) spins[v] = -1;
    int np = 0;
    while (np != N) {
        if (used >= nu) return -1;
        int v = (int)(u[used++] * Ns);
        if (spins[v] < 0) { spins[v] = 0; ++np; }
    }
    for (int v = 0; v < Ns; ++v)
        if (spins[v] < 0) spins[v] = 1;
    return (int)used;
}

/* State::InitFock (State.cpp:29-41): particle labels in increasing site order. */
void or_labels(int L, const int32_t *spins, int32_t *fock, int32_t *part)
{
    int Ns = L * L, N = Ns / 2;
    int cnt[2] = {0, 0};
    for (int v = 0; v < Ns; ++v) {
        int fl = spins[v];
        fock[fl * Ns + v] = cnt[fl];
        part[fl * N + cnt[fl]] = v;
        fock[(1 - fl) * Ns + v] = N;
        cnt[fl]++;
    }
}

/* SlaterDeterminant::Init (SlaterDeterminant.cpp:60-92): A(f, r) = Phi(orbital of p=f,
 * site of lattice particle r), DetInv per flavour; the amplitude is zero if either
 * determinant is singular.  Returns 1 when the amplitude is non-zero (m_amp_ok).
 * logabs/phase describe the product of both determinants. */
int or_init_walker(int L, const double *occin, const int32_t *spins, int32_t *fock,
                   int32_t *part, double *Ai, double *logabs, double *phase)
{
    int Ns = L * L, N = Ns / 2;
    const cplx *occ = (const cplx *)occin;
    or_labels(L, spins, fock, part);
    cplx *A = malloc(sizeof(cplx) * (size_t)N * N);
    double la = 0, ph = 0;
    int ok = 1;
    for (int fl = 0; fl < 2; ++fl) {
        for (int r = 0; r < N; ++r)
            for (int f = 0; f < N; ++f)
                A[(size_t)r * N + f] = occ[((size_t)fl * Ns + part[fl * N + r]) * N + f];
        double l1, p1;
        int sing = or_detinv(N, (const double *)A, Ai + 2 * (size_t)fl * N * N, &l1, &p1);
        if (sing) ok = 0;
        la += l1;
        ph += p1;
    }
    free(A);
    *logabs = la;
    *phase = remainder(ph, 2 * M_PI);
    return ok;
}

static int flavour_at(const int32_t *fock, int Ns, int N, int v)
{
    return fock[v] != N ? 0 : 1;
}

/* ProjHeis::measure for the ground state (ProjHeis.cpp:40-166): for each edge in order,
 * an antiparallel pair contributes -J(1/4 + 1/2 q_up q_dn), a parallel pair +J/4, J=1;
 * the result is divided by Lx*Ly (Val(0,0)).  q_fl is the single-hop ratio of the
 * particle of flavour fl on one end moving to the other end (LatticeStepper convention:
 * flavour fi hops vi->vj, flavour fj hops vj->vi). */
void or_eloc(int L, const double *occin, const int32_t *edges, const int32_t *fock,
             const int32_t *part, const double *Aiin, double *eloc)
{
    (void)part;
    int Ns = L * L, N = Ns / 2;
    const cplx *occ = (const cplx *)occin;
    const cplx *Ai = (const cplx *)Aiin;
    cplx h = 0;
    for (int e = 0; e < 2 * Ns; ++e) {
        int vi = edges[2 * e], vj = edges[2 * e + 1];
        int fi = flavour_at(fock, Ns, N, vi), fj = flavour_at(fock, Ns, N, vj);
        if (fi != fj) {
            cplx q[2];
            int from[2], to[2];
            from[fi] = vi; to[fi] = vj;
            from[fj] = vj; to[fj] = vi;
            for (int fl = 0; fl < 2; ++fl) {
                int ci = fock[fl * Ns + from[fl]];
                const cplx *A = Ai + (size_t)fl * N * N;
                const cplx *col = occ + ((size_t)fl * Ns + to[fl]) * N;
                cplx s = 0;
                for (int p = 0; p < N; ++p) s += A[ci + (size_t)p * N] * col[p];
                q[fl] = s;
            }
            h -= 0.25 + 0.5 * q[0] * q[1];
        } else {
            h += 0.25;
        }
    }
    h /= (double)Ns;
    eloc[0] = creal(h);
    eloc[1] = cimag(h);
}

/* StagMagnZ::measure (StagMagnZ.cpp:12-43): (1/Ns) sum_v exp(i pi (x+y)) (2 up_v - 1). */
void or_stagmagnz(int L, const int32_t *fock, double *mz)
{
    int Ns = L * L, N = Ns / 2;
    cplx m = 0;
    for (int v = 0; v < Ns; ++v) {
        int x = v % L, y = v / L;
        cplx ph = cexp((I * M_PI) * (double)(x + y));
        int up = fock[v] != N;
        m += ph * (double)(2 * up - 1);
    }
    m /= (double)Ns;
    mz[0] = creal(m);
    mz[1] = cimag(m);
}

/* StatSpinStruct::measure for two flavours (StatSpinStruct.cpp:16-135): adds one
 * measurement into acc[5][Ns][Ns] complex (szz, spm, smp, spp, smm). */
void or_statspinstruct(int L, const double *occin, const int32_t *fock, const int32_t *part,
                       const double *Aiin, double *accin)
{
    (void)part;
    int Ns = L * L, N = Ns / 2;
    const cplx *occ = (const cplx *)occin;
    const cplx *Ai = (const cplx *)Aiin;
    cplx *acc = (cplx *)accin;
    size_t S = (size_t)Ns * Ns;
    cplx *szz = acc, *spm = acc + S, *smp = acc + 2 * S;
    for (int vi = 0; vi < Ns; ++vi) {
        for (int vj = vi; vj < Ns; ++vj) {
            int fi = flavour_at(fock, Ns, N, vi), fj = flavour_at(fock, Ns, N, vj);
            if (fi == fj) {
                if (vi == vj) {
                    if (fi == 0) spm[vi * Ns + vj] += 1.0; else smp[vi * Ns + vj] += 1.0;
                    szz[vi * Ns + vj] += 0.25;
                } else {
                    szz[vi * Ns + vj] += 0.25;
                    szz[vj * Ns + vi] += 0.25;
                }
                continue;
            }
            cplx q[2];
            int from[2], to[2];
            from[fi] = vi; to[fi] = vj;
            from[fj] = vj; to[fj] = vi;
            for (int fl = 0; fl < 2; ++fl) {
                int ci = fock[fl * Ns + from[fl]];
                const cplx *A = Ai + (size_t)fl * N * N;
                const cplx *col = occ + ((size_t)fl * Ns + to[fl]) * N;
                cplx s = 0;
                for (int p = 0; p < N; ++p) s += A[ci + (size_t)p * N] * col[p];
                q[fl] = s;
            }
            cplx R = q[0] * q[1];
            szz[vi * Ns + vj] += -0.25;
            szz[vj * Ns + vi] += -0.25;
            if (fi == 0) {  /* up at vi, down at vj */
                smp[vi * Ns + vj] += -R;
                spm[vj * Ns + vi] += -R;
            } else {
                spm[vi * Ns + vj] += -R;
                smp[vj * Ns + vi] += -R;
            }
        }
    }
}

/* One Metropolis chain (MetroMC::Walk/Step, MetroMC.cpp:45-103) with LatticeStepper
 * proposals (LatticeStepper.cpp:36-147) on the ground-state Slater determinant, fed the
 * uniform stream u (4 per step).  Per step:
 *   u1 -> v1 = floor(Ns u1); u2 -> v2 = neighbour floor(4 u2) of v1; u3 -> internal state
 *   (one per site, unused); same flavour on both ends -> no-op proposal of ratio 1;
 *   otherwise each flavour's particle hops to the other end, ratio = |q_up q_dn|^2;
 *   u4 -> accept iff (ratio > 1 and u4 < 1) or u4 < ratio (MetroMC::YesNo :27-43).
 * After the step, if meas > 0 and (meas == 1 or s % meas == 0), measure E_loc, m_z and
 * the weight ln|amp|^2 into eloc/mz/logw (slot s / meas).  noop/accept/ratio get one
 * entry per step (ratio 1 for no-ops); cond[s] (may be NULL) is the sum over both flavours
 * of the condition number sum_p |Ai[ci,p]| |col_p| / |q| of the ratio's dot product (0 for
 * no-ops), which bounds how far two correct FP64 evaluations of the ratio may differ.
 * Returns the number of measurements. */
int or_walk(int L, const double *occin, const int32_t *nbr, int32_t *fock, int32_t *part,
            double *Aiin, double *logabs, double *phase, const double *u, int64_t nsteps,
            int64_t meas, int8_t *noop, int8_t *accept, double *ratio, double *eloc,
            double *mz, double *logw, double *cond)
{
    int Ns = L * L, N = Ns / 2;
    const cplx *occ = (const cplx *)occin;
    cplx *Ai = (cplx *)Aiin;
    int32_t *edges = malloc(sizeof(int32_t) * 4 * (size_t)Ns);
    or_edges(L, edges);
    int nm = 0;
    for (int64_t s = 0; s < nsteps; ++s) {
        const double *us = u + 4 * s;
        int v1 = (int)(Ns * us[0]);
        int v2 = nbr[4 * v1 + (int)(4 * us[1])];
        int f1 = flavour_at(fock, Ns, N, v1), f2 = flavour_at(fock, Ns, N, v2);
        int isnoop = (f1 == f2);
        double r = 1.0, kappa = 0.0;
        cplx q[2] = {1.0, 1.0};
        int from[2], to[2];
        if (!isnoop) {
            from[f1] = v1; to[f1] = v2;
            from[f2] = v2; to[f2] = v1;
            for (int fl = 0; fl < 2; ++fl) {
                int ci = fock[fl * Ns + from[fl]];
                const cplx *A = Ai + (size_t)fl * N * N;
                const cplx *col = occ + ((size_t)fl * Ns + to[fl]) * N;
                cplx sacc = 0;
                double mag = 0;
                for (int p = 0; p < N; ++p) {
                    sacc += A[ci + (size_t)p * N] * col[p];
                    mag += cabs(A[ci + (size_t)p * N]) * cabs(col[p]);
                }
                q[fl] = sacc;
                kappa += mag / cabs(sacc);
            }
            double a0 = cabs(q[0]), a1 = cabs(q[1]);
            r = (a0 * a0) * (a1 * a1);
        }
        if (cond) cond[s] = kappa;
        int acc = r > 1.0 ? (us[3] < 1.0) : (us[3] < r);
        if (noop) noop[s] = (int8_t)isnoop;
        if (accept) accept[s] = (int8_t)acc;
        if (ratio) ratio[s] = r;
        if (acc && !isnoop) {
            for (int fl = 0; fl < 2; ++fl) {
                int ci = fock[fl * Ns + from[fl]];
                or_sm_update(N, (double *)(Ai + (size_t)fl * N * N), ci,
                             (const double *)(occ + ((size_t)fl * Ns + to[fl]) * N));
                /* State::Hop (State.cpp:73-84) */
                fock[fl * Ns + from[fl]] = N;
                fock[fl * Ns + to[fl]] = ci;
                part[fl * N + ci] = to[fl];
            }
            cplx qq = q[0] * q[1];
            *logabs += log(cabs(qq));
            *phase = remainder(*phase + carg(qq), 2 * M_PI);
        }
        if (meas > 0 && (meas == 1 || s % meas == 0)) {
            if (eloc) or_eloc(L, occin, edges, fock, part, Aiin, eloc + 2 * nm);
            if (mz) or_stagmagnz(L, fock, mz + 2 * nm);
            if (logw) logw[nm] = 2 * (*logabs);
            ++nm;
        }
    }
    free(edges);
    return nm;
}

/* ---------------------------------------------------------------- channels
 * Amplitude of channel state k at a configuration, restated from
 * SlaterDeterminant::VirtUpdate (SlaterDeterminant.cpp:96-268) with the wavefunction
 * sign GetSign()*HopSign() (:133-137): the product over flavours of
 * det[Phi_fl(o_i, site of particle j)], rows o_i = the state's orbitals in increasing
 * order, columns = the flavour's particles by label.  Computed from scratch by LU with
 * partial pivoting, in log form (ln|a|, arg a), so it carries no update history.  The
 * reference evaluates the crossing terms of its capacitance determinant with the matrix
 * A of the last Init (linalg.cpp:196,320,457), which differs from these exact amplitudes
 * once particles have moved; these are the exact ones. */
static void logdet(int n, cplx *M, double *la, double *ph)
{
    double l = 0.0, p = 0.0;
    for (int k = 0; k < n; ++k) {
        int piv = k;
        double best = -1.0;
        for (int i = k; i < n; ++i) {
            double m = fabs(creal(M[i * n + k])) + fabs(cimag(M[i * n + k]));
            if (m > best) { best = m; piv = i; }
        }
        if (best == 0.0) { *la = -INFINITY; *ph = 0.0; return; }
        if (piv != k) {
            for (int j = 0; j < n; ++j) {
                cplx t = M[k * n + j]; M[k * n + j] = M[piv * n + j]; M[piv * n + j] = t;
            }
            p += M_PI;
        }
        cplx d = M[k * n + k];
        l += log(cabs(d));
        p += carg(d);
        for (int i = k + 1; i < n; ++i) {
            cplx f = M[i * n + k] / d;
            if (f == 0) continue;
            for (int j = k + 1; j < n; ++j) M[i * n + j] -= f * M[k * n + j];
        }
    }
    *la = l;
    *ph = remainder(p, 2 * M_PI);
}

/* tab[fl][orbital][site]; orbs[k][nup + ndn] (up orbitals then down, each increasing);
 * part[nup + ndn] = sites of the up particles by label, then of the down particles. */
void or_exc_amps(int Ns, int nup, int ndn, const double *tabin, int nk, const int32_t *orbs,
                 const int32_t *part, double *la, double *ph)
{
    const cplx *tab = (const cplx *)tabin;
    int nmax = nup > ndn ? nup : ndn;
    cplx *M = malloc(sizeof(cplx) * (size_t)nmax * nmax);
    for (int k = 0; k < nk; ++k) {
        double lt = 0.0, pt = 0.0;
        for (int fl = 0; fl < 2; ++fl) {
            int n = fl ? ndn : nup;
            const int32_t *o = orbs + (size_t)k * (nup + ndn) + (fl ? nup : 0);
            const int32_t *s = part + (fl ? nup : 0);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                    M[i * n + j] = tab[((size_t)fl * Ns + o[i]) * Ns + s[j]];
            double l, p;
            logdet(n, M, &l, &p);
            lt += l;
            pt += p;
        }
        la[k] = lt;
        ph[k] = remainder(pt, 2 * M_PI);
    }
    free(M);
}

static double logsumexp2(int nk, const double *la)
{
    double m = -INFINITY, s = 0.0;
    for (int k = 0; k < nk; ++k) if (la[k] > m) m = la[k];
    if (m == -INFINITY) return -INFINITY;
    for (int k = 0; k < nk; ++k) s += exp(2.0 * (la[k] - m));
    return 2.0 * m + log(s);
}

/* ProjHeis::measure on a channel (ProjHeis.cpp:40-140) at one configuration: amps a_k
 * (scaled by a common real factor so that max |a_k| = 1) and
 * h_k = sum over edges [antiparallel: -(a_k/4 + a_k(swapped)/2); parallel: +a_k/4]
 * (J = 1, Lattice.h:29; the swapped configuration keeps particle labels).  Returns the
 * scaled weight sum_k |a_k|^2; H and O are invariant under the common scale. */
double or_exc_local(int L, int nup, int ndn, const double *tabin, int nk, const int32_t *orbs,
                    const int32_t *site, const int32_t *part, double *ain, double *hin)
{
    int Ns = L * L, np = nup + ndn;
    cplx *a = (cplx *)ain, *h = (cplx *)hin;
    double *la = malloc(sizeof(double) * nk), *ph = malloc(sizeof(double) * nk);
    int32_t *pt = malloc(sizeof(int32_t) * np);
    int32_t *edges = malloc(sizeof(int32_t) * 4 * (size_t)Ns);
    or_edges(L, edges);
    or_exc_amps(Ns, nup, ndn, tabin, nk, orbs, part, la, ph);
    double m = -INFINITY, wn = 0.0;
    for (int k = 0; k < nk; ++k) if (la[k] > m) m = la[k];
    for (int k = 0; k < nk; ++k) {
        a[k] = exp(la[k] - m) * cexp(I * ph[k]);
        wn += creal(a[k] * conj(a[k]));
        h[k] = 0;
    }
    for (int e = 0; e < 2 * Ns; ++e) {
        int vi = edges[2 * e], vj = edges[2 * e + 1];
        int fi = site[vi] >> 16, fj = site[vj] >> 16;
        if (fi == fj) {
            for (int k = 0; k < nk; ++k) h[k] += 0.25 * a[k];
            continue;
        }
        int pi = (site[vi] & 0xFFFF) + (fi ? nup : 0), pj = (site[vj] & 0xFFFF) + (fj ? nup : 0);
        memcpy(pt, part, sizeof(int32_t) * np);
        pt[pi] = vj;
        pt[pj] = vi;
        or_exc_amps(Ns, nup, ndn, tabin, nk, orbs, pt, la, ph);
        for (int k = 0; k < nk; ++k)
            h[k] -= 0.25 * a[k] + 0.5 * exp(la[k] - m) * cexp(I * ph[k]);
    }
    free(la); free(ph); free(pt); free(edges);
    return wn;
}

/* One chain on a channel (MetroMC::Walk/Step, MetroMC.cpp:45-103; LatticeStepper::trystep
 * and step, LatticeStepper.cpp:36-147): the weight is w = sum_k |a_k|^2 (trystep's
 * BigDouble sum of norm(qs), :90-91), proposals and the accept rule are those of or_walk,
 * 4 uniforms per step.  site[v] = (flavour << 16) | label, part as in or_exc_amps.
 * After step s, when meas > 0 and (meas == 1 or s % meas == 0), ProjHeis::measure
 * (ProjHeis.cpp:40-154) adds, with amps a_k at the configuration and
 * h_k = sum_bonds [antiparallel: -(a_k/4 + a_k(swapped)/2); parallel: a_k/4],
 *   HO[k1][k2] += conj(a_k1) h_k2 / w   and   HO[nk + k1][k2] += conj(a_k1) a_k2 / w,
 * and logw[m] = ln w (OverlapTrack).  noop/accept/ratio get one entry per step.
 * Returns the number of measurements. */
int or_exc_walk(int L, int nup, int ndn, const double *tabin, int nk, const int32_t *orbs,
                const int32_t *nbr, int32_t *site, int32_t *part, const double *u,
                int64_t nsteps, int64_t meas, int8_t *noop, int8_t *accept, double *ratio,
                double *HOin, double *logw)
{
    int Ns = L * L, np = nup + ndn;
    cplx *HO = (cplx *)HOin;
    double *la = malloc(sizeof(double) * nk), *ph = malloc(sizeof(double) * nk);
    double *lb = malloc(sizeof(double) * nk), *pb = malloc(sizeof(double) * nk);
    cplx *a = malloc(sizeof(cplx) * nk), *h = malloc(sizeof(cplx) * nk);
    int32_t *pt = malloc(sizeof(int32_t) * np);
    or_exc_amps(Ns, nup, ndn, tabin, nk, orbs, part, la, ph);
    double lw = logsumexp2(nk, la);
    int nm = 0;
    for (int64_t s = 0; s < nsteps; ++s) {
        const double *us = u + 4 * s;
        int v1 = (int)(Ns * us[0]);
        if (v1 >= Ns) v1 = Ns - 1;
        int ed = (int)(4 * us[1]);
        if (ed > 3) ed = 3;
        int v2 = nbr[4 * v1 + ed];
        int f1 = site[v1] >> 16, f2 = site[v2] >> 16;
        double r = 1.0;
        int acc;
        if (f1 == f2) {
            acc = us[3] < 1.0;
            if (noop) noop[s] = 1;
        } else {
            if (noop) noop[s] = 0;
            int p1 = (site[v1] & 0xFFFF) + (f1 ? nup : 0), p2 = (site[v2] & 0xFFFF) + (f2 ? nup : 0);
            memcpy(pt, part, sizeof(int32_t) * np);
            pt[p1] = v2;
            pt[p2] = v1;
            or_exc_amps(Ns, nup, ndn, tabin, nk, orbs, pt, lb, pb);
            double lw2 = logsumexp2(nk, lb);
            r = exp(lw2 - lw);
            acc = r > 1.0 ? (us[3] < 1.0) : (us[3] < r);
            if (acc) {
                memcpy(part, pt, sizeof(int32_t) * np);
                int t = site[v1]; site[v1] = site[v2]; site[v2] = t;
                memcpy(la, lb, sizeof(double) * nk);
                memcpy(ph, pb, sizeof(double) * nk);
                lw = lw2;
            }
        }
        if (accept) accept[s] = (int8_t)acc;
        if (ratio) ratio[s] = r;
        if (meas > 0 && (meas == 1 || s % meas == 0)) {
            double wn = or_exc_local(L, nup, ndn, tabin, nk, orbs, site, part, (double *)a, (double *)h);
            for (int k1 = 0; k1 < nk; ++k1)
                for (int k2 = 0; k2 < nk; ++k2) {
                    HO[(size_t)k1 * nk + k2] += conj(a[k1]) * h[k2] / wn;
                    HO[(size_t)(nk + k1) * nk + k2] += conj(a[k1]) * a[k2] / wn;
                }
            if (logw) logw[nm] = lw;
            ++nm;
        }
    }
    free(la); free(ph); free(lb); free(pb); free(a); free(h); free(pt);
    return nm;
}
