/* gpvmc oracle: plain-C FP64 restatement of the reference's Metropolis VMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or the timed CPU
 * baseline).  The product path (libgpvmc_hip.so) never links or calls it.
 *
 * Conventions (all follow the reference; file:line citations in gpvmc_oracle.c):
 *  - sites v = y*L + x; 4 neighbours per site in the reference's edge order;
 *  - flavour 0 = "up" (sigma=+1 in the orbitals), flavour 1 = "down";
 *  - complex arrays are interleaved (re, im) doubles;
 *  - orbital table phi[fl][f][r] = Phi_fl(orbital f, site r), f,r < Ns;
 *  - occupied table occ[fl][r][p] = Phi_fl(orbital of wavefunction particle p, site r);
 *  - inverse matrices use the reference's memory layout: Ai[fl][r + p*N] is
 *    (A^-1)(lattice particle r, wavefunction particle p), column-major;
 *  - fock[fl][site] = particle label or N (empty); part[fl][label] = site.
 */
#ifndef GPVMC_ORACLE_H
#define GPVMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void or_neighbours(int L, int32_t *nbr);
void or_edges(int L, int32_t *edges);
int  or_mbz(int L, double bcx, double bcy, int32_t *fock2qn);
int  or_orbitals(int L, double phi, double neel, double neel_exp, double bcx, double bcy,
                 double *table);
void or_ground_occ(int L, const double *table, double *occ);

int  or_detinv(int N, const double *A, double *Ai, double *logabs, double *phase);
void or_ratio(int N, const double *Ai, int ci, const double *col, double *q);
void or_sm_update(int N, double *Ai, int ci, const double *col);

int  or_raninit(int L, const double *u, int64_t nu, int32_t *spins);
void or_labels(int L, const int32_t *spins, int32_t *fock, int32_t *part);
int  or_init_walker(int L, const double *occ, const int32_t *spins, int32_t *fock,
                    int32_t *part, double *Ai, double *logabs, double *phase);

int  or_walk(int L, const double *occ, const int32_t *nbr, int32_t *fock, int32_t *part,
             double *Ai, double *logabs, double *phase, const double *u, int64_t nsteps,
             int64_t meas, int8_t *noop, int8_t *accept, double *ratio, double *eloc,
             double *mz, double *logw, double *cond);

void or_eloc(int L, const double *occ, const int32_t *edges, const int32_t *fock,
             const int32_t *part, const double *Ai, double *eloc);
void or_stagmagnz(int L, const int32_t *fock, double *mz);
void or_statspinstruct(int L, const double *occ, const int32_t *fock, const int32_t *part,
                       const double *Ai, double *acc);

/* channels (exciton states): exact amplitudes by LU; a chain on w = sum_k |a_k|^2 */
void or_exc_amps(int Ns, int nup, int ndn, const double *tab, int nk, const int32_t *orbs,
                 const int32_t *part, double *la, double *ph);
double or_exc_local(int L, int nup, int ndn, const double *tab, int nk, const int32_t *orbs,
                    const int32_t *site, const int32_t *part, double *amps, double *heis);
int  or_exc_walk(int L, int nup, int ndn, const double *tab, int nk, const int32_t *orbs,
                 const int32_t *nbr, int32_t *site, int32_t *part, const double *u,
                 int64_t nsteps, int64_t meas, int8_t *noop, int8_t *accept, double *ratio,
                 double *HO, double *logw);

#ifdef __cplusplus
}
#endif
#endif
