/* gpvmc_host.hpp - C++ host classes over libgpvmc_hip.so with the reference's names.
 *
 * The reference builds one Markov chain per process from a StagFlux* wavefunction, a
 * SlaterDeterminant amplitude, a LatticeStepper and MetroMC with Quantity estimators
 * (vmc.cpp:140-317).  These classes keep that object graph and those names, for a batch of
 * chains on one GPU: a SlaterDeterminant here is the amplitude of every walker of a batch
 * (one gpvmc_ctx), MetroMC::Walk steps all of them, and Quantity::save reads the device's
 * FP64 sums.  Errors throw gpvmc::Error (the reference's -DEXCEPT behaviour) instead of
 * aborting.
 *
 *   StagFluxWaveFunction   StagFluxWaveFunction.cpp:8-163 (orbital table), with the states
 *                          of StagFluxGroundState.cpp:3-15, StagFluxTransExciton.cpp:7-44,
 *                          StagFluxLongExciton.cpp:6-66 and the hop graph of
 *                          WaveFunction::AddState / State::HopSign (WaveFunction.cpp:39-67,
 *                          State.cpp:86-115)
 *   SlaterDeterminant      SlaterDeterminant::Init / Amp (SlaterDeterminant.cpp:60-92)
 *   LatticeStepper         LatticeStepper (LatticeStepper.cpp:36-147), proposal and step on
 *                          the device
 *   Quantity, ProjHeis, StagMagnZ, StatSpinStruct, OverlapTrack
 *                          Quantity::init / save (Quantity.cpp:19-27, ScalarQuantity.cpp,
 *                          MatrixQuantity.cpp)
 *   MetroMC                MetroMC::Walk / AddQuantity / Rejection (MetroMC.cpp:27-103)
 */
#ifndef GPVMC_HOST_HPP
#define GPVMC_HOST_HPP

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gpvmc.h"

namespace gpvmc {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

/* Orbitals of the staggered-flux + Neel mean-field state and the states of a channel.
 * phi is in units of pi, as vmc's --phi (vmc.cpp:120); bc = 1 is antiperiodic. */
class StagFluxWaveFunction {
  public:
    enum Channel { GroundState, TransExciton, LongExciton };
    StagFluxWaveFunction(int L, double phi, double neel, double neel_exp = 0.0, double bcx = 1.0,
                         double bcy = 1.0, Channel channel = GroundState, int qx = 0, int qy = 0);

    int GetL() const { return L_; }
    Channel GetChannel() const { return ch_; }
    int GetNExc() const { return nk_; }
    int GetNpt(int fl) const { return npt_[fl]; }
    /* Phi[fl][orbital][site] (WaveFunction::MatEl, WaveFunction.h:61-66) */
    std::complex<double> MatEl(int orbital, int site, int fl) const
    {
        return table_[((size_t)fl * Ns_ + orbital) * Ns_ + site];
    }
    const std::vector<std::complex<double>> &Table() const { return table_; }
    /* occupied orbitals of state k, n_up of flavour 0 then n_dn of flavour 1, increasing */
    const std::vector<int32_t> &State(int k) const { return states_[k]; }
    /* hop path from state c to state k in flavour fl: (from, to) orbital pairs */
    std::vector<std::pair<int32_t, int32_t>> GetHop(int c, int k, int fl) const;
    int HopSign(int c, int k) const { return hopsign_[(size_t)c * nk_ + k]; }
    /* the graph as the C ABI takes it (pointers into this object) */
    gpvmc_channel GraphABI() const;

  private:
    int L_, Ns_, N_, nk_;
    int npt_[2];
    Channel ch_;
    std::vector<std::complex<double>> table_;
    std::vector<std::vector<int32_t>> states_;
    std::vector<int32_t> state_orb_, hop_n_, hop_orb_, hopsign_;
};

/* Initial configurations: LatticeState::RanInit's role (LatticeState.cpp:37-82), a
 * Fisher-Yates shuffle per walker drawn from Philox4x32-10 (key seed, counter
 * (i, 0, w, 1)); n_up particles of flavour 0.  spins[w][site] in {0, 1}. */
std::vector<int32_t> RanInit(int L, int n_walkers, int n_up, uint64_t seed);

/* log-magnitude and phase of an amplitude (the reference's BigComplex m_amp) */
struct LogAmp {
    double logabs, phase;
};

class SlaterDeterminant {
  public:
    /* precision: GPVMC_FP64 or GPVMC_MIXED */
    SlaterDeterminant(const StagFluxWaveFunction &wav, int n_walkers, int precision, uint64_t seed,
                      int device = 0);
    ~SlaterDeterminant();
    SlaterDeterminant(const SlaterDeterminant &) = delete;
    SlaterDeterminant &operator=(const SlaterDeterminant &) = delete;

    /* SlaterDeterminant::Init on every walker's configuration; returns m_amp_ok per walker */
    std::vector<int32_t> Init(const std::vector<int32_t> &spins);
    /* FP64 recomputation of the inverses at the current configurations */
    void Refresh();
    /* Amp() of every walker */
    std::vector<LogAmp> Amp() const;
    /* current configurations spins[w][site] */
    std::vector<int32_t> Spins() const;
    const StagFluxWaveFunction &GetWaveFunction() const { return wav_; }
    int NWalkers() const { return W_; }
    gpvmc_ctx *Context() const { return ctx_; }

  private:
    const StagFluxWaveFunction &wav_;
    int W_;
    gpvmc_ctx *ctx_ = nullptr;
};

class LatticeStepper {
  public:
    explicit LatticeStepper(SlaterDeterminant &amp) : amp_(amp) {}
    SlaterDeterminant &GetAmp() const { return amp_; }

  private:
    SlaterDeterminant &amp_;
};

/* Quantity: init() clears the sums, save() reads them (per walker, averaged over its
 * measurements) */
class Quantity {
  public:
    virtual ~Quantity() = default;
    virtual uint32_t Mask() const = 0;
    virtual void init(SlaterDeterminant &amp);
    virtual void save(SlaterDeterminant &amp) = 0;
    const std::string &GetName() const { return name_; }

  protected:
    explicit Quantity(std::string name) : name_(std::move(name)) {}
    std::string name_;
};

/* ProjHeis: ground state, E_loc per site averaged per walker (Val(0,0), ProjHeis.cpp:140-145);
 * channels, the H and O matrices averaged per walker (ProjHeis.cpp:146-153) */
class ProjHeis : public Quantity {
  public:
    ProjHeis() : Quantity("ProjHeis") {}
    uint32_t Mask() const override { return GPVMC_MEAS_PROJHEIS; }
    void save(SlaterDeterminant &amp) override;
    std::vector<std::complex<double>> energy;  // [W] (ground state)
    std::vector<std::complex<double>> H, O;    // [W][nk][nk] (channels)
    int nk = 0;
};

/* StagMagnZ: staggered magnetisation m_z averaged per walker (StagMagnZ.cpp:12-43) */
class StagMagnZ : public Quantity {
  public:
    StagMagnZ() : Quantity("StagMagnZ") {}
    uint32_t Mask() const override { return GPVMC_MEAS_STAGMAGNZ; }
    void save(SlaterDeterminant &amp) override;
    std::vector<std::complex<double>> mz;  // [W]
    std::vector<double> mz2;               // [W] <|m_z|^2>
};

/* StatSpinStruct: the 5 x Ns x Ns correlators summed over the walkers (StatSpinStruct.cpp:16-135) */
class StatSpinStruct : public Quantity {
  public:
    StatSpinStruct() : Quantity("StatSpinStruct") {}
    uint32_t Mask() const override { return GPVMC_MEAS_STATSPIN; }
    void init(SlaterDeterminant &amp) override;
    void save(SlaterDeterminant &amp) override;
    std::vector<std::complex<double>> corr;  // [5][Ns][Ns], averaged over all measurements
};

/* OverlapTrack: the mean of ln w per walker (OverlapTrack.cpp:6-29) */
class OverlapTrack : public Quantity {
  public:
    OverlapTrack() : Quantity("OverlapTrack") {}
    uint32_t Mask() const override { return GPVMC_MEAS_OVERLAP; }
    void save(SlaterDeterminant &amp) override;
    std::vector<double> mean_logw;  // [W]
};

class MetroMC {
  public:
    explicit MetroMC(LatticeStepper &stepper) : st_(stepper) {}
    void AddQuantity(Quantity *q) { quantities_.push_back(q); }
    /* MetroMC::Walk(len, meas): len steps of every walker, measuring every meas steps
     * (meas = 0: no measurement) */
    void Walk(int64_t len, int64_t meas);
    /* rejected / proposed over every walker since the quantities' init */
    double Rejection() const;

  private:
    LatticeStepper &st_;
    std::vector<Quantity *> quantities_;
};

}  // namespace gpvmc

#endif
