// vmc_host: a batch of chains through the C++ host classes, as vmc.cpp wires one chain
// (vmc.cpp:140-317): wavefunction, SlaterDeterminant, LatticeStepper, MetroMC with its
// quantities, thermalisation then measured walks.
//
// Two forms:
//   vmc_host --L=8 --samples=10 --meas_projheis ...   the reference's vmc flag set and defaults
//       (vmc.cpp:69-94, ArgParse's --key=value / --flag syntax) and its sample / save loop
//       (vmc.cpp:295-317); every walker is one chain (the reference's rank).  Extra flags:
//       --walkers (default 1) and --precision=fp64|mixed (default fp64).  Each quantity's
//       records go to <dir>/<prefix>-<Name>.h5 (/rank-<r>/data-<k> with rej, time and
//       statistics, FileManager.cpp:465-551), the prefix from --prefix or, when it is
//       negative, from <dir>/.runfile (FileManager.cpp:20-64); <prefix>-WaveFunction.h5 at
//       the start (WaveFunction.cpp:85-137) and <prefix>-FockState.h5 at the end
//       (State.cpp:117-166); --spinstate=<FockState file> starts from its configurations
//       (vmc.cpp:178-220).
//   vmc_host L walkers phi neel channel(0 ground, 1 trans, 2 long) qx qy therm_sweeps
//       samples meas_interval precision(0 FP64, 1 mixed) seed
//       one measured walk; prints one JSON line of per-walker means.
#include <chrono>
#include <ctime>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <exception>
#include <fstream>
#include <map>
#include <string>

#include "gpvmc_host.hpp"

namespace {

struct Flags {
    std::map<std::string, bool> bo;
    std::map<std::string, double> num;  // size_t, int and double flags
    std::map<std::string, std::string> str;
};

Flags default_flags()
{
    Flags f;
    f.num = {{"L", 8},
             {"samples", 10},
             {"samples_saves", 1},
             {"samples_saves_stat", 100},
             {"qx", 0},
             {"qy", 0},
             {"meas_interv", 0},
             {"therm", 100},
             {"prefix", -1},
             {"verbose", 1},
             {"seed", (double)std::time(nullptr)},
             {"phi", 0.085},
             {"neel", 0.055},
             {"neel_exp", 0.0},
             {"phase_shift_x", 1.0},
             {"phase_shift_y", 1.0},
             {"jr", 0.0},
             {"walkers", 1}};
    f.bo = {{"track_stagmagn", false}, {"track_overlap", false},       {"meas_projheis", false},
            {"meas_stagmagn", false},  {"meas_statspinstruct", false}, {"Neel_init", false}};
    f.str = {{"dir", "."}, {"spinstate", ""}, {"channel", "groundstate"}, {"precision", "fp64"}};
    return f;
}

void print_help(const Flags &f)
{
    std::printf("Available options:\n");
    for (const auto &a : f.num) std::printf("--%s=<number> (default %g)\n", a.first.c_str(), a.second);
    for (const auto &a : f.bo) std::printf("--%s\n", a.first.c_str());
    for (const auto &a : f.str) std::printf("--%s=<string> (default \"%s\")\n", a.first.c_str(), a.second.c_str());
}

// --key=value, --key value, or --flag (booleans); returns 1 for --help
int parse(int argc, char **argv, Flags &f)
{
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--help" || a == "-h") return 1;
        if (a.rfind("--", 0) != 0) throw gpvmc::Error("unexpected argument " + a);
        a = a.substr(2);
        std::string key = a, val;
        const size_t eq = a.find('=');
        bool has_val = eq != std::string::npos;
        if (has_val) {
            key = a.substr(0, eq);
            val = a.substr(eq + 1);
        }
        if (f.bo.count(key)) {
            if (has_val) throw gpvmc::Error("flag --" + key + " takes no value");
            f.bo[key] = true;
            continue;
        }
        if (!f.num.count(key) && !f.str.count(key)) throw gpvmc::Error("unknown option --" + key);
        if (!has_val) {
            if (i + 1 >= argc) throw gpvmc::Error("option --" + key + " needs a value");
            val = argv[++i];
        }
        if (f.num.count(key)) {
            char *end = nullptr;
            const double v = std::strtod(val.c_str(), &end);
            if (end == val.c_str() || *end) throw gpvmc::Error("option --" + key + ": not a number: " + val);
            f.num[key] = v;
        } else {
            f.str[key] = val;
        }
    }
    return 0;
}

// FileManager's run number: --prefix when >= 0, else one more than <dir>/.runfile (0 when
// absent), written back (FileManager.cpp:20-64)
int run_number(const std::string &dir, int prefix)
{
    if (prefix >= 0) return prefix;
    int num = 0;
    std::ifstream in(dir + "/.runfile");
    if (in >> num)
        ++num;
    else
        num = 0;
    std::ofstream out(dir + "/.runfile");
    if (!(out << num << "\n")) throw gpvmc::Error("error writing runfile in " + dir);
    return num;
}

int run_flags(int argc, char **argv)
{
    Flags f = default_flags();
    if (parse(argc, argv, f)) {
        print_help(f);
        return 0;
    }
    const int L = (int)f.num["L"], W = (int)f.num["walkers"];
    const std::string channel = f.str["channel"];
    gpvmc::StagFluxWaveFunction::Channel ch;
    if (channel == "groundstate")
        ch = gpvmc::StagFluxWaveFunction::GroundState;
    else if (channel == "trans")
        ch = gpvmc::StagFluxWaveFunction::TransExciton;
    else if (channel == "long")
        ch = gpvmc::StagFluxWaveFunction::LongExciton;
    else
        throw gpvmc::Error("unknown channel " + channel);
    int prec;
    if (f.str["precision"] == "fp64")
        prec = GPVMC_FP64;
    else if (f.str["precision"] == "mixed")
        prec = GPVMC_MIXED;
    else
        throw gpvmc::Error("precision must be fp64 or mixed");
    if (f.num["jr"] != 0.0) throw gpvmc::Error("--jr != 0 is not supported (the reference's vmc does not use it)");
    if (f.bo["Neel_init"] && ch == gpvmc::StagFluxWaveFunction::TransExciton)
        throw gpvmc::Error("--Neel_init fills L*L/2 sites per flavour; the trans channel has L*L/2 +- 1");
    const size_t samples = (size_t)f.num["samples"], saves = (size_t)f.num["samples_saves"],
                 stat = (size_t)f.num["samples_saves_stat"], therm = (size_t)f.num["therm"];
    size_t meas = (size_t)f.num["meas_interv"];
    if (!meas) meas = (size_t)L * L;
    const uint64_t seed = (uint64_t)f.num["seed"];
    std::printf("seed=%llu\n", (unsigned long long)seed);

    // FileManager and its root attributes (vmc.cpp:101-117)
    gpvmc::FileManager fm(W);
    const int num = run_number(f.str["dir"], (int)f.num["prefix"]);
    fm.SetOutput(f.str["dir"], std::to_string(num));
    fm.StatPerSample() = (int)saves;
    for (const auto &a : f.bo) fm.FileAttribute(a.first, a.second ? 1.0 : 0.0);
    for (const auto &a : f.num) fm.FileAttribute(a.first, a.second);
    for (const auto &a : f.str) fm.FileAttribute(a.first, a.second);
    fm.FileAttribute("stagflux_wav", 1.0);
    fm.FileAttribute("sfpnxphz_wav", 0.0);
    fm.FileAttribute("sfpnzphx_wav", 0.0);
    fm.FileAttribute("gitversion", "gpvmc_amd-" + std::to_string(gpvmc_version()));

    gpvmc::StagFluxWaveFunction wav(L, f.num["phi"], f.num["neel"], f.num["neel_exp"], f.num["phase_shift_x"],
                                    f.num["phase_shift_y"], ch, (int)f.num["qx"], (int)f.num["qy"]);
    wav.Save(fm);
    gpvmc::SlaterDeterminant amp(wav, W, prec, seed);
    std::vector<int32_t> spins, ok;
    if (!f.str["spinstate"].empty()) {
        // vmc.cpp:178-220: the configurations of a FockState file, labels in site order
        spins = gpvmc::LoadFockState(f.str["spinstate"], L, W);
        for (int w = 0; w < W; ++w) {
            int nu = 0;
            for (int v = 0; v < L * L; ++v) nu += spins[(size_t)w * L * L + v] == 0;
            if (nu != wav.GetNpt(0))
                throw gpvmc::Error("--spinstate: chain " + std::to_string(w) + " has " + std::to_string(nu) +
                                   " particles of flavour 0, the channel " + std::to_string(wav.GetNpt(0)));
        }
        ok = amp.Init(spins);
        for (int x : ok)
            if (!x) throw gpvmc::Error("--spinstate: a configuration of the file has a zero amplitude");
    } else if (f.bo["Neel_init"]) {
        // vmc.cpp:163-177: flavour 0 on the sites with x + y odd
        spins.resize((size_t)W * L * L);
        for (int w = 0; w < W; ++w)
            for (int y = 0; y < L; ++y)
                for (int x = 0; x < L; ++x) spins[(size_t)w * L * L + y * L + x] = (x + y) % 2 ? 0 : 1;
        ok = amp.Init(spins);
        for (int x : ok)
            if (!x) throw gpvmc::Error("the Neel configuration has a zero amplitude");
    } else {
        // vmc.cpp:222-241: up to 10 RanInit draws per chain, drawn and initialised on the
        // device (draw t on Philox attempt t, the draws of gpvmc::RanInit)
        ok = amp.InitRandom(seed, 10);
        for (int x : ok)
            if (!x) throw gpvmc::Error("could not find a non-singular starting configuration.");
    }
    gpvmc::LatticeStepper stepper(amp);
    gpvmc::MetroMC mc(stepper);
    gpvmc::ProjHeis heis(&fm);
    gpvmc::StagMagnZ magn(&fm);
    gpvmc::StatSpinStruct sss(&fm);
    gpvmc::StagMagnTrack mtrack(&fm);
    gpvmc::OverlapTrack otrack(&fm);
    const bool ground = ch == gpvmc::StagFluxWaveFunction::GroundState;
    if (f.bo["meas_projheis"]) mc.AddQuantity(&heis);
    if (ground && f.bo["meas_stagmagn"]) mc.AddQuantity(&magn);
    if (ground && f.bo["meas_statspinstruct"]) mc.AddQuantity(&sss);
    if (f.bo["track_stagmagn"]) mc.AddQuantity(&mtrack);
    if (f.bo["track_overlap"]) mc.AddQuantity(&otrack);
    if (therm) mc.Walk((int64_t)(therm * L * L), 0);
    double walk_s = 0;
    for (size_t sample = 0; sample < samples; ++sample)
        for (size_t m = 0; m < saves; ++m) {
            const auto t0 = std::chrono::steady_clock::now();
            mc.Walk((int64_t)(meas * stat), (int64_t)meas);
            walk_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            for (gpvmc::Quantity *q : mc.GetQuantities()) q->save(amp);
            fm.DataAttribute("rej", mc.RejectionPerChain());
            fm.DataAttribute("time", walk_s);
            fm.DataAttribute("statistics", (double)((m + 1) * stat));
            fm.Write();
            if (f.num["verbose"] > 0)
                std::printf("sample %zu save %zu: rejection %.6f\n", sample, m, mc.Rejection());
        }
    gpvmc::SaveFockState(fm, L, amp.Spins());  // State::Save at the end of the run (vmc.cpp:329)
    std::printf("{\"prefix\": %d, \"chains\": %d, \"rejection\": %.15g, \"quantities\": [", num, W, mc.Rejection());
    const std::vector<std::string> names = fm.Names();
    for (size_t i = 0; i < names.size(); ++i) std::printf("%s\"%s\"", i ? ", " : "", names[i].c_str());
    std::printf("]}\n");
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc > 1 && std::strncmp(argv[1], "-", 1) == 0) {
        try {
            return run_flags(argc, argv);
        } catch (const std::exception &e) {
            std::fprintf(stderr, "vmc_host: %s\n", e.what());
            return 1;
        }
    }
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
        // comes from the Fisher-Yates draw of seed + 1000 * (its try)
        std::vector<int32_t> spins = gpvmc::RandomSpins(L, W, wav.GetNpt(0), seed);
        std::vector<int32_t> ok = amp.Init(spins);
        for (int tries = 1; tries < 50; ++tries) {
            bool all = true;
            for (int x : ok) all = all && x;
            if (all) break;
            const std::vector<int32_t> fresh = gpvmc::RandomSpins(L, W, wav.GetNpt(0), seed + 1000ull * tries);
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
