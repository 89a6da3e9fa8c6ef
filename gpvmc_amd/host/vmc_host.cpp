// vmc_host: a batch of chains through the C++ host classes, as vmc.cpp wires one chain
// (vmc.cpp:140-317): wavefunction, SlaterDeterminant, LatticeStepper, MetroMC with
// ProjHeis / StagMagnZ / OverlapTrack, thermalisation then measured walks.  Prints one JSON
// line of per-walker means.
//
// usage: vmc_host L walkers phi neel channel(0 ground, 1 trans, 2 long) qx qy therm_sweeps
//        samples meas_interval precision(0 FP64, 1 mixed) seed
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <exception>

#include "gpvmc_host.hpp"

int main(int argc, char **argv)
{
    if (argc != 13) {
        std::fprintf(stderr, "usage: %s L walkers phi neel channel qx qy therm samples meas prec seed\n", argv[0]);
        return 2;
    }
    try {
        const int L = std::atoi(argv[1]), W = std::atoi(argv[2]);
        const double phi = std::atof(argv[3]), neel = std::atof(argv[4]);
        const int channel = std::atoi(argv[5]), qx = std::atoi(argv[6]), qy = std::atoi(argv[7]);
        const long therm = std::atol(argv[8]), samples = std::atol(argv[9]), meas = std::atol(argv[10]);
        const int prec = std::atoi(argv[11]);
        const unsigned long long seed = std::strtoull(argv[12], nullptr, 10);
        gpvmc::StagFluxWaveFunction wav(L, phi, neel, 0.0, 1.0, 1.0, (gpvmc::StagFluxWaveFunction::Channel)channel,
                                        qx, qy);
        gpvmc::SlaterDeterminant amp(wav, W, prec, seed);
        // singular starts are redrawn, as vmc.cpp:230-241 does: walker w's configuration
        // comes from RanInit with seed + 1000 * (its try)
        std::vector<int32_t> spins = gpvmc::RanInit(L, W, wav.GetNpt(0), seed);
        std::vector<int32_t> ok = amp.Init(spins);
        for (int tries = 1; tries < 50; ++tries) {
            bool all = true;
            for (int x : ok) all = all && x;
            if (all) break;
            const std::vector<int32_t> fresh = gpvmc::RanInit(L, W, wav.GetNpt(0), seed + 1000ull * tries);
            for (int w = 0; w < W; ++w)
                if (!ok[w]) std::copy(fresh.begin() + (size_t)w * L * L, fresh.begin() + (size_t)(w + 1) * L * L,
                                      spins.begin() + (size_t)w * L * L);
            ok = amp.Init(spins);
        }
        gpvmc::LatticeStepper stepper(amp);
        gpvmc::MetroMC mc(stepper);
        gpvmc::ProjHeis heis;
        gpvmc::StagMagnZ magn;
        gpvmc::OverlapTrack ovl;
        mc.AddQuantity(&heis);
        if (channel == 0) mc.AddQuantity(&magn);
        mc.AddQuantity(&ovl);
        mc.Walk(therm * L * L, 0);
        heis.init(amp);
        mc.Walk(samples * meas, meas);
        heis.save(amp);
        ovl.save(amp);
        int nok = 0;
        for (int x : ok) nok += x;
        std::printf("{\"L\": %d, \"walkers\": %d, \"init_ok\": %d, \"rejection\": %.15g", L, W, nok, mc.Rejection());
        if (channel == 0) {
            magn.save(amp);
            double e = 0, m2 = 0;
            for (int w = 0; w < W; ++w) {
                e += heis.energy[w].real();
                m2 += magn.mz2[w];
            }
            std::printf(", \"energy_per_site\": %.15g, \"stag_magn_sq\": %.15g", e / W, m2 / W);
        } else {
            // trace of H over trace of O averaged over walkers (a scalar summary of the matrices)
            double th = 0, to = 0;
            const int nk = heis.nk;
            for (int w = 0; w < W; ++w)
                for (int k = 0; k < nk; ++k) {
                    th += heis.H[((size_t)w * nk + k) * nk + k].real();
                    to += heis.O[((size_t)w * nk + k) * nk + k].real();
                }
            std::printf(", \"nk\": %d, \"trace_H\": %.15g, \"trace_O\": %.15g", nk, th / W, to / W);
        }
        double lw = 0;
        for (int w = 0; w < W; ++w) lw += ovl.mean_logw[w];
        // a checksum of the final configurations (sum over walkers and sites of
        // spin * (site + 1) * (walker + 1))
        const std::vector<int32_t> sp = amp.Spins();
        long long cs = 0;
        for (int w = 0; w < W; ++w)
            for (int v = 0; v < L * L; ++v) cs += (long long)sp[(size_t)w * L * L + v] * (v + 1) * (w + 1);
        std::printf(", \"mean_logw\": %.15g, \"spins_checksum\": %lld}\n", lw / W, cs);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "vmc_host: %s\n", e.what());
        return 1;
    }
}
