// Test-infrastructure harness that drives the reference's own classes, compiled from
// the reference sources where they lie (see oracle/Makefile, target `ref`).  It is the
// checker's checker: its output becomes golden fixtures under tests/golden/ that pin
// the oracle restatement.  Nothing in the product links or loads it.
//
// Build notes (no stand-ins are written for anything the image lacks):
//  * RanGen is built with -DUSE_RNG_STD (RanGen.cpp:46-47,114-115): uniform() is
//    std::rand()/RAND_MAX, so the stream is reproducible from the seed through glibc's
//    rand() on any machine with the same libc.
//  * zdotu_sub.f (the Fortran shim over BLAS zdotu) is compiled with the LLVM flang of
//    the ROCm toolchain.  flang returns COMPLEX functions in registers, as the GNU
//    convention does, so the harness selects MKL's GNU interface layer before any BLAS call
//    (MKL's default layer returns them through a hidden argument).
//  * FileManager, ProjHeis and the Quantity classes are the reference's own; HDF5 is the
//    image's (/opt/conda).
//
// Modes (all output is one JSON object on stdout):
//   lattice L
//   orbitals L phi_in_units_of_pi neel neel_exp bcx bcy
//   walk L phi_in_units_of_pi neel neel_exp seed nsteps snap_every
//   excgraph L channel(trans|long) qx qy
//   time L phi_in_units_of_pi neel seed nsteps   (wall time of nsteps MetroMC::Step)
//   h5layout dir L phi_in_units_of_pi neel seed   (files written by FileManager, WaveFunction::Save
//                                                and State::Save; the data written on stdout)
//   excwalk channel L qx qy phi_in_units_of_pi neel seed nsteps meas
//                                               (a channel chain with ProjHeis every meas steps)
//   eloc L phi_in_units_of_pi neel seed nsteps meas   (ground-state ProjHeis E_loc along a chain)
#include <cstdio>
#include <ctime>
#include <cstdlib>
#include <cmath>
#include <complex>
#include <vector>
#include <string>
#include <iostream>
#include <map>
#include <algorithm>
#include <numeric>
#include <stdexcept>
// the reference headers below expose only what vmc.cpp needs; the harness also reads
// the inverse matrices and running weights, which are private/protected members.
#define private public
#define protected public
#include "SquareLattice.h"
#include "LatticeState.h"
#include "StagFluxGroundState.h"
#include "StagFluxTransExciton.h"
#include "StagFluxLongExciton.h"
#include "SlaterDeterminant.h"
#include "LatticeStepper.h"
#include "MetroMC.h"
#include "RanGen.h"
#include "ProjHeis.h"
#include "FileManager.h"
#undef private
#undef protected
#include <hdf5.h>

// MKL's interface-layer switch (mkl_rt): LP64 integers, GNU convention for complex returns
extern "C" int MKL_Set_Interface_Layer(int);
static const int kMklLp64Gnu = 0 + 2;

using namespace std;

static void pr_bigc(const BigComplex& b)
{
    printf("[%.17g,%.17g,%d]", real(b.m_c), imag(b.m_c), b.m_e);
}

static void pr_bigd(const BigDouble& b)
{
    printf("[%.17g,%d]", b.m_c, b.m_e);
}

static void pr_fock(const LatticeState* sp)
{
    // flavour index per site: 0 (up) or 1 (down)
    printf("[");
    for(size_t v=0; v<sp->GetNsites(); ++v){
        int fl = (sp->Getfs()[0][v] != sp->GetNpt()[0]) ? 0 : 1;
        printf("%s%d", v ? "," : "", fl);
    }
    printf("]");
}

// Number of values the glibc stream of `seed` has given so far (the RanGen of a USE_RNG_STD
// build draws std::rand()): draw the next value, re-seed, count the draws up to its first
// occurrence, then leave the stream where it was.  A 31-bit value recurring within the few
// thousand draws of a start-up is the only way to miscount; a caller checks the chain.
static long stream_position(int seed)
{
    int next = std::rand();
    std::srand(seed);
    long n = 0;
    while(std::rand() != next) ++n;
    std::srand(seed);
    for(long i=0; i<n; ++i) std::rand();
    return n;
}

static void pr_labels(const LatticeState* sp, size_t fl)
{
    printf("[");
    for(size_t p=0; p<sp->GetNpt()[fl]; ++p)
        printf("%s%zu", p ? "," : "", sp->Getpt()[fl][p]);
    printf("]");
}

static int mode_lattice(int L)
{
    SquareLattice slat(L, L);
    printf("{\"L\":%d,\"neighbours\":[", L);
    for(size_t v=0; v<slat.GetNv(); ++v){
        const Vertex* vx = slat.GetVertices()[v];
        printf("%s[", v ? "," : "");
        for(size_t e=0; e<vx->edges.size(); ++e)
            printf("%s%zu", e ? "," : "", vx->edges[e]->GetOther(vx)->idx);
        printf("]");
    }
    printf("],\"edges\":[");
    for(size_t e=0; e<slat.GetEdges().size(); ++e)
        printf("%s[%zu,%zu]", e ? "," : "", slat.GetEdges()[e]->first->idx,
               slat.GetEdges()[e]->second->idx);
    printf("]}\n");
    return 0;
}

static int mode_orbitals(int L, double phi, double neel, double ne, double bcx, double bcy)
{
    StagFluxGroundState wav(nullptr, L, L, phi * M_PI, neel, ne, {bcx, bcy});
    size_t Ns = L * L;
    printf("{\"L\":%d,\"phi\":%.17g,\"neel\":%.17g,\"neel_exp\":%.17g,\"bc\":[%.17g,%.17g],",
           L, phi, neel, ne, bcx, bcy);
    printf("\"re\":[");
    for(size_t fl=0; fl<2; ++fl)
        for(size_t f=0; f<Ns; ++f)
            for(size_t r=0; r<Ns; ++r)
                printf("%s%.17g", (fl||f||r) ? "," : "", real(wav.MatEl(f, r, fl)));
    printf("],\"im\":[");
    for(size_t fl=0; fl<2; ++fl)
        for(size_t f=0; f<Ns; ++f)
            for(size_t r=0; r<Ns; ++r)
                printf("%s%.17g", (fl||f||r) ? "," : "", imag(wav.MatEl(f, r, fl)));
    printf("],\"fock2qn\":[");
    for(size_t f=0; f<Ns; ++f)
        printf("%s[%zu,%zu,%zu]", f ? "," : "", wav.m_fock2qn[3*f], wav.m_fock2qn[3*f+1],
               wav.m_fock2qn[3*f+2]);
    printf("]}\n");
    return 0;
}

// vmc.cpp:221-259 (random start with retries, then the kmax reference state), then
// MetroMC::Step nsteps times (MetroMC.cpp:90-103), recording per step the predicted
// weight ratio and whether the lattice state moved.
static int mode_walk(int L, double phi, double neel, double ne, int seed, long nsteps,
                     long snap_every)
{
    size_t N = L * L / 2;
    RanGen::srand(seed);
    SquareLattice slat(L, L);
    LatticeState sp(nullptr, &slat, {N, N}, {1, 1});
    StagFluxGroundState wav(nullptr, L, L, phi * M_PI, neel, ne, {1.0, 1.0});
    SlaterDeterminant amp(&sp, &wav);
    vector<vector<size_t>> pop = {{N}, {N}};
    int tc = 0;
    vector<vector<int>> inits;
    while(amp.Amp() == 0.0 && tc < 10){
        sp.RanInit(pop);
        amp.Init();
        vector<int> f(sp.GetNsites());
        for(size_t v=0; v<sp.GetNsites(); ++v)
            f[v] = (sp.Getfs()[0][v] != sp.GetNpt()[0]) ? 0 : 1;
        inits.push_back(f);
        tc++;
    }
    if(tc == 10){ fprintf(stderr, "no regular start\n"); return 1; }
    BigDouble maxq = norm(amp.Amp());
    size_t kmax = wav.GetCurrentStateIndex();
    for(size_t k=0; k<wav.GetNExc(); ++k){
        wav.Hop(k);
        amp.Init();
        if(norm(amp.Amp()) > maxq){ kmax = k; maxq = norm(amp.Amp()); }
    }
    wav.Hop(kmax);
    amp.Init();
    LatticeStepper step(&sp, &wav, &amp);
    MetroMC varmc(&step, nullptr);

    printf("{\"L\":%d,\"phi\":%.17g,\"neel\":%.17g,\"neel_exp\":%.17g,\"seed\":%d,"
           "\"nsteps\":%ld,", L, phi, neel, ne, seed, nsteps);
    printf("\"init_attempts\":[");
    for(size_t t=0; t<inits.size(); ++t){
        printf("%s[", t ? "," : "");
        for(size_t v=0; v<inits[t].size(); ++v) printf("%s%d", v ? "," : "", inits[t][v]);
        printf("]");
    }
    printf("],\"init_fock\":");
    pr_fock(&sp);
    printf(",\"init_labels\":[");
    pr_labels(&sp, 0); printf(","); pr_labels(&sp, 1);
    printf("],\"init_amp\":");
    pr_bigc(amp.m_amp);
    printf(",\"init_weight\":");
    pr_bigd(varmc.m_state_weight);
    // Snapshot of the initial inverse matrices (column-major Ai(r,f) at r+f*N)
    printf(",\"init_inv\":[");
    for(size_t fl=0; fl<2; ++fl)
        for(size_t i=0; i<N*N; ++i)
            printf("%s%.17g,%.17g", (fl||i) ? "," : "", real(amp.m_mati[fl][i]),
                   imag(amp.m_mati[fl][i]));
    printf("]");

    vector<double> ratio(nsteps);
    vector<int> moved(nsteps);
    vector<vector<int>> snaps;
    vector<BigDouble> snapw;
    vector<BigComplex> snapa;
    for(long s=0; s<nsteps; ++s){
        vector<size_t> before = sp.Getfs()[0];
        BigDouble w0 = varmc.m_state_weight;
        varmc.Step(false);
        ratio[s] = double(step.m_prev_weight / w0);
        moved[s] = (before != sp.Getfs()[0]);
        if(snap_every > 0 && (s + 1) % snap_every == 0){
            vector<int> f(sp.GetNsites());
            for(size_t v=0; v<sp.GetNsites(); ++v)
                f[v] = (sp.Getfs()[0][v] != sp.GetNpt()[0]) ? 0 : 1;
            snaps.push_back(f);
            snapw.push_back(step.weight());
            snapa.push_back(amp.m_amp);
        }
    }
    printf(",\"ratio\":[");
    for(long s=0; s<nsteps; ++s) printf("%s%.17g", s ? "," : "", ratio[s]);
    printf("],\"moved\":[");
    for(long s=0; s<nsteps; ++s) printf("%s%d", s ? "," : "", moved[s]);
    printf("],\"snap_every\":%ld,\"snaps\":[", snap_every);
    for(size_t t=0; t<snaps.size(); ++t){
        printf("%s[", t ? "," : "");
        for(size_t v=0; v<snaps[t].size(); ++v) printf("%s%d", v ? "," : "", snaps[t][v]);
        printf("]");
    }
    printf("],\"snap_weight\":[");
    for(size_t t=0; t<snapw.size(); ++t){ if(t) printf(","); pr_bigd(snapw[t]); }
    printf("],\"snap_amp\":[");
    for(size_t t=0; t<snapa.size(); ++t){ if(t) printf(","); pr_bigc(snapa[t]); }
    printf("],\"final_labels\":[");
    pr_labels(&sp, 0); printf(","); pr_labels(&sp, 1);
    printf("],\"final_inv\":[");
    for(size_t fl=0; fl<2; ++fl)
        for(size_t i=0; i<N*N; ++i)
            printf("%s%.17g,%.17g", (fl||i) ? "," : "", real(amp.m_mati[fl][i]),
                   imag(amp.m_mati[fl][i]));
    printf("]}\n");
    return 0;
}

// Excitation graph of an exciton channel (WaveFunction.cpp:132-160 via the channel
// constructors), and the fermionic sign of each hop path from state 0 (State.cpp:86-115).
static int mode_excgraph(int L, const string& channel, size_t qx, size_t qy)
{
    WaveFunction* wav = nullptr;
    if(channel == "trans")
        wav = new StagFluxTransExciton(nullptr, L, L, 0.1 * M_PI, 0.05, 0.0, {1.0, 1.0}, {qx, qy});
    else if(channel == "long")
        wav = new StagFluxLongExciton(nullptr, L, L, 0.1 * M_PI, 0.05, 0.0, {1.0, 1.0}, {qx, qy});
    else
        wav = new StagFluxGroundState(nullptr, L, L, 0.1 * M_PI, 0.05, 0.0, {1.0, 1.0});
    size_t nexc = wav->GetNExc();
    printf("{\"L\":%d,\"channel\":\"%s\",\"q\":[%zu,%zu],\"nexc\":%zu,\"npt\":[%zu,%zu],",
           L, channel.c_str(), qx, qy, nexc, wav->GetNpt()[0], wav->GetNpt()[1]);
    // occupied orbitals of the current (first) state, by wavefunction particle label
    printf("\"state0_part\":[");
    for(size_t fl=0; fl<2; ++fl){
        printf("%s[", fl ? "," : "");
        for(size_t p=0; p<wav->GetNpt()[fl]; ++p)
            printf("%s%zu", p ? "," : "", wav->Getpt()[fl][p]);
        printf("]");
    }
    printf("],\"hops\":[");
    for(size_t k=0; k<nexc; ++k){
        const vector<hop_path_t>& h = wav->GetHop(k);
        printf("%s[", k ? "," : "");
        for(size_t fl=0; fl<2; ++fl){
            printf("%s[", fl ? "," : "");
            for(size_t i=0; i<h[fl].size(); ++i)
                printf("%s[%zu,%zu]", i ? "," : "", h[fl][i].first, h[fl][i].second);
            printf("]");
        }
        printf("]");
    }
    printf("],\"hopsign\":[");
    for(size_t k=0; k<nexc; ++k)
        printf("%s%d", k ? "," : "", wav->HopSign(wav->GetHop(k)));
    // the full all-pairs path matrix m_exc[i][j] for the first few states
    size_t nshow = min<size_t>(nexc, 6);
    printf("],\"exc_block\":[");
    for(size_t i=0; i<nshow; ++i){
        printf("%s[", i ? "," : "");
        for(size_t j=0; j<nshow; ++j){
            printf("%s[", j ? "," : "");
            for(size_t fl=0; fl<2; ++fl){
                printf("%s[", fl ? "," : "");
                const hop_path_t& h = static_cast<WaveFunction*>(wav)->m_exc[i][j][fl];
                for(size_t t=0; t<h.size(); ++t)
                    printf("%s[%zu,%zu]", t ? "," : "", h[t].first, h[t].second);
                printf("]");
            }
            printf("]");
        }
        printf("]");
    }
    printf("]}\n");
    delete wav;
    return 0;
}

// Times the reference's own hot loop: nsteps of MetroMC::Step on one chain after the
// vmc.cpp:221-263 start-up (random start with retries).
static int mode_time(int L, double phi, double neel, int seed, long nsteps)
{
    size_t N = L * L / 2;
    RanGen::srand(seed);
    SquareLattice slat(L, L);
    LatticeState sp(nullptr, &slat, {N, N}, {1, 1});
    StagFluxGroundState wav(nullptr, L, L, phi * M_PI, neel, 0.0, {1.0, 1.0});
    SlaterDeterminant amp(&sp, &wav);
    vector<vector<size_t>> pop = {{N}, {N}};
    int tc = 0;
    while(amp.Amp() == 0.0 && tc < 10){ sp.RanInit(pop); amp.Init(); tc++; }
    if(tc == 10){ fprintf(stderr, "no regular start\n"); return 1; }
    LatticeStepper step(&sp, &wav, &amp);
    MetroMC varmc(&step, nullptr);
    timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for(long s=0; s<nsteps; ++s) varmc.Step(true);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double sec = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    printf("{\"L\":%d,\"nsteps\":%ld,\"seconds\":%.6f,\"steps_per_s\":%.3f,\"rejection\":%.6f}\n",
           L, nsteps, sec, nsteps / sec, varmc.Rejection());
    return 0;
}

// FileManager::Write / WriteSimple, WaveFunction::Save and State::Save into `dir` (prefix 7
// for the ground state at (L, phi, neel), prefix 8 for the transverse exciton at q = (1, 2)):
// RanInit from the glibc stream of `seed` (LatticeState.cpp:37-82), a ProjHeis-named
// stream written 4 times with 2 saves per sample and the rej/time/statistics attributes.
static int mode_h5layout(const string& dir, int L, double phi, double neel, int seed)
{
    size_t N = L * L / 2;
    RanGen::srand(seed);
    SquareLattice slat(L, L);
    {
        FileManager fm(dir, 7);
        fm.StatPerSample() = 2;
        fm.FileAttribute("L", double(L));
        fm.FileAttribute("phi", phi);
        fm.FileAttribute("channel", string("groundstate"));
        StagFluxGroundState wav(&fm, L, L, phi * M_PI, neel, 0.0, {1.0, 1.0});
        wav.Save();
        LatticeState sp(&fm, &slat, {N, N}, {1, 1});
        vector<vector<size_t>> pop = {{N}, {N}};
        sp.RanInit(pop);
        printf("{\"fock\":");
        pr_fock(&sp);
        printf(",\"saves\":[");
        for(int s=0; s<4; ++s){
            double m[4] = {0.25 * s - 1.0, 1.0 / (s + 3.0), -0.5 * s, 2.0 + s};
            fm.FileStream("ProjHeis").Write(2, 2, m);
            fm.DataAttribute("rej", 0.1 * s);
            fm.DataAttribute("time", 0.0);
            fm.DataAttribute("statistics", double((s % 2 + 1) * 100));
            fm.Write();
            printf("%s[%.17g,%.17g,%.17g,%.17g]", s ? "," : "", m[0], m[1], m[2], m[3]);
        }
        printf("]");
        sp.Save();
    }
    {
        FileManager fm(dir, 8);
        fm.FileAttribute("L", double(L));
        fm.FileAttribute("channel", string("trans"));
        StagFluxTransExciton wav(&fm, L, L, phi * M_PI, neel, 0.0, {1.0, 1.0}, {1, 2});
        wav.Save();
    }
    printf("}\n");
    return 0;
}

// vmc.cpp:221-259 start-up (RanInit retries, then the kmax state) on a channel, then
// MetroMC::Step nsteps times with ProjHeis::measure after step s when s % meas == 0
// (MetroMC.cpp:45-103, ProjHeis.cpp:40-153): per-step predicted weight ratios, moves, the
// starting configuration and the H/O sums.  The reference's own DetUpdate/InvUpdate run,
// crossing terms from the A of the last Init included (linalg.cpp:196,320,457).
static int mode_excwalk(const string& channel, int L, size_t qx, size_t qy, double phi, double neel, int seed,
                        long nsteps, long meas)
{
    size_t N = L * L / 2;
    RanGen::srand(seed);
    SquareLattice slat(L, L);
    vector<size_t> npt = {N, N};
    if(channel == "trans") npt = {N + 1, N - 1};
    LatticeState sp(nullptr, &slat, npt, {1, 1});
    WaveFunction* wav = nullptr;
    if(channel == "trans")
        wav = new StagFluxTransExciton(nullptr, L, L, phi * M_PI, neel, 0.0, {1.0, 1.0}, {qx, qy});
    else
        wav = new StagFluxLongExciton(nullptr, L, L, phi * M_PI, neel, 0.0, {1.0, 1.0}, {qx, qy});
    SlaterDeterminant amp(&sp, wav);
    vector<vector<size_t>> pop = {{npt[0]}, {npt[1]}};
    int tc = 0;
    while(amp.Amp() == 0.0 && tc < 10){ sp.RanInit(pop); amp.Init(); tc++; }
    if(tc == 10){ fprintf(stderr, "no regular start\n"); return 1; }
    BigDouble maxq = norm(amp.Amp());
    size_t kmax = wav->GetCurrentStateIndex();
    for(size_t k=0; k<wav->GetNExc(); ++k){
        wav->Hop(k);
        amp.Init();
        if(norm(amp.Amp()) > maxq){ kmax = k; maxq = norm(amp.Amp()); }
    }
    wav->Hop(kmax);
    amp.Init();
    LatticeStepper step(&sp, wav, &amp);
    MetroMC varmc(&step, nullptr);
    ProjHeis heis(&step, nullptr, &slat);
    heis.init();
    const long pos = stream_position(seed);
    printf("{\"channel\":\"%s\",\"L\":%d,\"q\":[%zu,%zu],\"seed\":%d,\"nsteps\":%ld,\"meas\":%ld,"
           "\"init_attempts\":%d,\"stream_pos\":%ld,\"kmax\":%zu,\"init_fock\":", channel.c_str(), L, qx, qy,
           seed, nsteps, meas, tc, pos, kmax);
    pr_fock(&sp);
    vector<double> ratio(nsteps);
    vector<int> moved(nsteps);
    for(long s=0; s<nsteps; ++s){
        vector<size_t> before = sp.Getfs()[0];
        BigDouble w0 = varmc.m_state_weight;
        varmc.Step(true);
        ratio[s] = double(step.m_prev_weight / w0);
        moved[s] = (before != sp.Getfs()[0]);
        if(meas > 0 && s % meas == 0) heis.measure();
    }
    printf(",\"ratio\":[");
    for(long s=0; s<nsteps; ++s) printf("%s%.17g", s ? "," : "", ratio[s]);
    printf("],\"moved\":[");
    for(long s=0; s<nsteps; ++s) printf("%s%d", s ? "," : "", moved[s]);
    printf("],\"n_meas\":%zu,\"ho\":[", heis.m_meas);
    for(size_t i=0; i<heis.m_vals.size(); ++i)
        printf("%s%.17g,%.17g", i ? "," : "", real(heis.m_vals[i]), imag(heis.m_vals[i]));
    printf("],\"final_fock\":");
    pr_fock(&sp);
    printf("}\n");
    delete wav;
    return 0;
}

// ground-state ProjHeis E_loc per site along a chain (vmc.cpp start-up, MetroMC::Step, and
// ProjHeis::measure after step s when s % meas == 0): the per-measurement values and the
// configurations they were measured at
static int mode_eloc(int L, double phi, double neel, int seed, long nsteps, long meas)
{
    size_t N = L * L / 2;
    RanGen::srand(seed);
    SquareLattice slat(L, L);
    LatticeState sp(nullptr, &slat, {N, N}, {1, 1});
    StagFluxGroundState wav(nullptr, L, L, phi * M_PI, neel, 0.0, {1.0, 1.0});
    SlaterDeterminant amp(&sp, &wav);
    vector<vector<size_t>> pop = {{N}, {N}};
    int tc = 0;
    while(amp.Amp() == 0.0 && tc < 10){ sp.RanInit(pop); amp.Init(); tc++; }
    if(tc == 10){ fprintf(stderr, "no regular start\n"); return 1; }
    LatticeStepper step(&sp, &wav, &amp);
    MetroMC varmc(&step, nullptr);
    ProjHeis heis(&step, nullptr, &slat);
    printf("{\"L\":%d,\"seed\":%d,\"nsteps\":%ld,\"meas\":%ld,\"eloc\":[", L, seed, nsteps, meas);
    vector<vector<int>> confs;
    int m = 0;
    for(long s=0; s<nsteps; ++s){
        varmc.Step(true);
        if(meas > 0 && s % meas == 0){
            heis.init();
            heis.measure();
            printf("%s[%.17g,%.17g]", m++ ? "," : "", real(heis.m_vals[0]), imag(heis.m_vals[0]));
            vector<int> f(sp.GetNsites());
            for(size_t v=0; v<sp.GetNsites(); ++v) f[v] = (sp.Getfs()[0][v] != sp.GetNpt()[0]) ? 0 : 1;
            confs.push_back(f);
        }
    }
    printf("],\"confs\":[");
    for(size_t t=0; t<confs.size(); ++t){
        printf("%s[", t ? "," : "");
        for(size_t v=0; v<confs[t].size(); ++v) printf("%s%d", v ? "," : "", confs[t][v]);
        printf("]");
    }
    printf("]}\n");
    return 0;
}

int main(int argc, char** argv)
{
    if(argc < 2){ fprintf(stderr, "usage: see header\n"); return 2; }
    MKL_Set_Interface_Layer(kMklLp64Gnu);
    string mode = argv[1];
    if(mode == "h5layout" && argc == 7)
        return mode_h5layout(argv[2], atoi(argv[3]), atof(argv[4]), atof(argv[5]), atoi(argv[6]));
    if(mode == "excwalk" && argc == 11)
        return mode_excwalk(argv[2], atoi(argv[3]), (size_t)atoi(argv[4]), (size_t)atoi(argv[5]), atof(argv[6]),
                            atof(argv[7]), atoi(argv[8]), atol(argv[9]), atol(argv[10]));
    if(mode == "eloc" && argc == 8)
        return mode_eloc(atoi(argv[2]), atof(argv[3]), atof(argv[4]), atoi(argv[5]), atol(argv[6]), atol(argv[7]));
    if(mode == "lattice" && argc == 3) return mode_lattice(atoi(argv[2]));
    if(mode == "orbitals" && argc == 8)
        return mode_orbitals(atoi(argv[2]), atof(argv[3]), atof(argv[4]), atof(argv[5]),
                             atof(argv[6]), atof(argv[7]));
    if(mode == "walk" && argc == 9)
        return mode_walk(atoi(argv[2]), atof(argv[3]), atof(argv[4]), atof(argv[5]),
                         atoi(argv[6]), atol(argv[7]), atol(argv[8]));
    if(mode == "time" && argc == 7)
        return mode_time(atoi(argv[2]), atof(argv[3]), atof(argv[4]), atoi(argv[5]), atol(argv[6]));
    if(mode == "excgraph" && argc == 6)
        return mode_excgraph(atoi(argv[2]), argv[3], (size_t)atoi(argv[4]), (size_t)atoi(argv[5]));
    fprintf(stderr, "bad arguments\n");
    return 2;
}
