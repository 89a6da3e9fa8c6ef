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
 *   Quantity, ProjHeis, StagMagnZ, StatSpinStruct, OverlapTrack, StagMagnTrack
 *                          Quantity::init / save (Quantity.cpp:19-27, ScalarQuantity.cpp,
 *                          VectorQuantity.cpp, MatrixQuantity.cpp, OverlapTrack.cpp,
 *                          StagMagnTrack.cpp)
 *   MetroMC                MetroMC::Walk / AddQuantity / Rejection (MetroMC.cpp:27-103)
 *   MatStream, FileManager the running mean over saves and the data-k bookkeeping
 *                          (FileManager.h:78-108, FileManager.cpp:465-551)
 */
#ifndef GPVMC_HOST_HPP
#define GPVMC_HOST_HPP

#include <complex>
#include <cstdint>
#include <functional>
#include <map>
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
    /* quantum numbers (kx, ky, band) of orbital f at [3 f .. 3 f + 2] (the fock2qn table,
     * StagFluxWaveFunction.cpp:24-41) */
    const std::vector<int32_t> &QuantumNumbers() const { return qn_; }
    /* WaveFunction::Save (WaveFunction.cpp:85-137): <dir>/<prefix>-WaveFunction.h5 with, per
     * flavour fl, the int dataset states_<fl>[state][orbital] = 1 where the channel state
     * occupies the orbital, and on it the unsigned attributes kx, ky, band[orbital] */
    void Save(const class FileManager &fm) const;

  private:
    int L_, Ns_, N_, nk_;
    int npt_[2];
    Channel ch_;
    std::vector<int32_t> qn_;
    std::vector<std::complex<double>> table_;
    std::vector<std::vector<int32_t>> states_;
    std::vector<int32_t> state_orb_, hop_n_, hop_orb_, hopsign_;
};

/* LatticeState::RanInit (LatticeState.cpp:37-82) for two flavours with one internal state
 * per site: flavour 0 takes the site v = floor(u Ns) of each uniform u until n_up distinct
 * sites are taken (a site already taken is drawn again), then flavour 1 fills every
 * remaining site.  Returns spins[site] in {0, 1}; *used (may be null) receives the number of
 * uniforms consumed. */
std::vector<int32_t> RanInitDraw(int L, int n_up, const std::function<double()> &uniform, int64_t *used = nullptr);
/* RanInitDraw for a batch: walker w's uniforms from Philox4x32-10, key seed, counter
 * (i, attempt, w, 2) for its i-th uniform, u = (x + 1/2) 2^-32 in (0, 1) (the open interval
 * of RanGen's uniform_pos, RanGen.cpp:86).  spins[w][site]. */
std::vector<int32_t> RanInit(int L, int n_walkers, int n_up, uint64_t seed, int attempt = 0);
/* Half-filled configurations by a Fisher-Yates shuffle per walker from Philox4x32-10 (key
 * seed, counter (i, 0, w, 1)): the Python host's model.random_spins. */
std::vector<int32_t> RandomSpins(int L, int n_walkers, int n_up, uint64_t seed);

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
    /* the random start of vmc.cpp:221-242 drawn on the device (gpvmc_init_random): RanInit
     * draws with redraws of zero-amplitude walkers, up to max_attempts draws; returns
     * m_amp_ok per walker and, when asked, the draws each used */
    std::vector<int32_t> InitRandom(uint64_t seed, int max_attempts, std::vector<int32_t> *attempts = nullptr);
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

/* MatStream::Write (FileManager.h:78-108).  The first Write after a restart (Stat() == 0)
 * stores the m x n doubles; every later Write must have the same shape (else Error) and
 * replaces the stored matrix with the running mean (stat * old + new) / (stat + 1). */
class MatStream {
  public:
    void Write(int m, int n, const double *mat);
    int &Stat() { return stat_; }
    int Rows() const { return nrow_; }
    int Cols() const { return ncol_; }
    const std::vector<double> &Mat() const { return mat_; }

  private:
    int nrow_ = 0, ncol_ = 0, stat_ = 0;
    std::vector<double> mat_;
};

/* FileManager's stream and dataset bookkeeping (FileManager.cpp:465-551) for a batch of
 * chains: the reference keeps one FileManager per rank (one chain), here chain c is walker c.
 * FileStream(name, c) is the MatStream a quantity's save writes into.  Write() stores every
 * stream's matrix as a dataset of its chain: a new one (data-k, k = the chain's count so far)
 * when the stream's Stat() is 1, else over the last one; a stream whose Stat() reaches
 * StatPerSample() restarts at its next Write (one dataset per sample of StatPerSample()
 * saves).  The data attributes set before a Write (rej, time, statistics) go onto every
 * dataset it writes; with SetOutput they also go to <dir>/<prefix>-<name>.h5. */
class FileManager {
  public:
    struct Dataset {
        int rows = 0, cols = 0;
        std::vector<double> mat;
        std::map<std::string, double> attr;
    };
    explicit FileManager(int n_chains);
    /* HDF5 output: every later Write() also stores the datasets it writes in
     * <dir>/<prefix>-<name>.h5, as the reference's FileManager::Write (FileManager.cpp:465-551):
     * a new file gets the root attributes, "type" = name and "date"; chain c's datasets go
     * to group RankGroup(c) as data-<k> (2-D doubles) with the data attributes. */
    void SetOutput(const std::string &dir, const std::string &prefix);
    const std::string &Dir() const { return dir_; }
    const std::string &Prefix() const { return prefix_; }
    /* FileManager::WriteSimple (FileManager.cpp:555-575): creates <dir>/<prefix>-<name>.h5
     * with the root attributes and "type"; returns the open HDF5 file id (the caller closes
     * it with H5Fclose) */
    int64_t WriteSimple(const std::string &name) const;
    /* group of chain c: "/rank-0" for a single chain (the reference without MPI), else
     * "/rank-<c + 1>" (the calculation ranks 1 .. W of an MPI run, rank 0 being its writer) */
    std::string RankGroup(int c) const;
    int NChains() const { return n_chains_; }
    MatStream &FileStream(const std::string &name, int chain);
    int &StatPerSample() { return stat_per_sample_; }
    void FileAttribute(const std::string &key, double v) { fattr_num_[key] = v; }
    void FileAttribute(const std::string &key, const std::string &v) { fattr_str_[key] = v; }
    /* one value for every chain, or one per chain */
    void DataAttribute(const std::string &key, double v);
    void DataAttribute(const std::string &key, const std::vector<double> &per_chain);
    void Write();
    std::vector<std::string> Names() const;
    const std::vector<Dataset> &Datasets(const std::string &name, int chain) const;

  private:
    void WriteHDF5(const std::string &name, const std::vector<std::vector<Dataset>> &rec,
                   const std::vector<int> &written) const;
    int n_chains_, stat_per_sample_ = 1;
    std::string dir_, prefix_;
    std::map<std::string, std::vector<MatStream>> streams_;
    std::map<std::string, std::vector<std::vector<Dataset>>> data_;
    std::map<std::string, std::vector<double>> dattr_;
    std::map<std::string, double> fattr_num_;
    std::map<std::string, std::string> fattr_str_;
};

/* State::Save (State.cpp:117-166): <dir>/<prefix>-FockState.h5 with, for chain c, the int
 * datasets RankGroup(c)/flavour-<fl>[site] = 1 where flavour fl occupies the site.
 * spins[c][site] in {0, 1}. */
void SaveFockState(const FileManager &fm, int L, const std::vector<int32_t> &spins);
/* The configurations of a FockState file for n_chains chains (--spinstate, vmc.cpp:178-220):
 * chain c reads RankGroup(c) of a FileManager with n_chains chains. */
std::vector<int32_t> LoadFockState(const std::string &path, int L, int n_chains);

/* The sample grouping of tools/concatenate.cpp:107-133: records (datasets) with the given
 * "statistics" are split into nsamp samples of about equal total statistics.  The record
 * indices are sorted by statistics (std::sort, so equal statistics keep the order the
 * reference's sort gives them); while more than 2 nsamp + 1 remain, the smallest and the
 * largest go together into sample n (n cycling 0 .. nsamp-1); each one left over, largest
 * first, goes into the first sample of smallest total so far.  Returns the members of each
 * sample (in the order they were added) and each sample's total. */
struct RecordBunches {
    std::vector<std::vector<int>> members;
    std::vector<int> statistics;
};
RecordBunches BunchRecords(int nsamp, const std::vector<int> &statistics);
/* tools/concatenate.cpp:26-233: merges the /rank-r/data-k records of several output files
 * of one quantity (ranks from 1 up, data sets from 0 up, as long as they exist; file order,
 * then rank, then k) into `out` (created; an existing file is an error): the root attributes
 * of the first file, and /rank-1/data-s = the plain mean of sample s's records (summed in
 * member order; the first member's datatype and shape) with the int attribute "statistics"
 * = the sample's total.  nsamp = 0 makes one sample per record.  Errors (no record, more
 * samples than records, records of different shapes) throw. */
void Concatenate(const std::vector<std::string> &in, const std::string &out, int nsamp);

/* per-measurement traces a quantity asks MetroMC::Walk to record */
enum Trace : uint32_t { TraceLogW = 1u, TraceStagMagn = 2u };

/* Quantity: init() clears the sums (MetroMC::Walk calls it before every measuring walk, as
 * MetroMC.cpp:49-51 does); save() averages them per walker over its measurements and, when
 * the quantity has a FileManager, writes each walker's value into FileStream(name, walker)
 * with the reference's shape. */
class Quantity {
  public:
    virtual ~Quantity() = default;
    virtual uint32_t Mask() const = 0;
    virtual uint32_t Traces() const { return 0; }
    /* the traces of one walk: logw[w][m] = ln w, mz[w][m] complex, m < n_meas */
    virtual void TakeTraces(int, int64_t, const double *, const double *) {}
    virtual void init(SlaterDeterminant &amp);
    virtual void save(SlaterDeterminant &amp) = 0;
    const std::string &GetName() const { return name_; }

  protected:
    Quantity(std::string name, FileManager *fm) : name_(std::move(name)), fm_(fm) {}
    std::string name_;
    FileManager *fm_;
};

/* ProjHeis: ground state, E_loc per site averaged per walker (Val(0,0), ProjHeis.cpp:140-145);
 * channels, the H and O matrices averaged per walker (ProjHeis.cpp:146-153).  Stream shape:
 * 2 nk x 2 nk doubles, rows H then O (ground state: [E_loc; 1]). */
class ProjHeis : public Quantity {
  public:
    explicit ProjHeis(FileManager *fm = nullptr) : Quantity("ProjHeis", fm) {}
    uint32_t Mask() const override { return GPVMC_MEAS_PROJHEIS; }
    void save(SlaterDeterminant &amp) override;
    std::vector<std::complex<double>> energy;  // [W] (ground state)
    std::vector<std::complex<double>> H, O;    // [W][nk][nk] (channels)
    int nk = 0;
};

/* StagMagnZ: staggered magnetisation m_z averaged per walker (StagMagnZ.cpp:12-43); stream
 * 1 x 2 */
class StagMagnZ : public Quantity {
  public:
    explicit StagMagnZ(FileManager *fm = nullptr) : Quantity("StagMagnZ", fm) {}
    uint32_t Mask() const override { return GPVMC_MEAS_STAGMAGNZ; }
    void save(SlaterDeterminant &amp) override;
    std::vector<std::complex<double>> mz;  // [W]
    std::vector<double> mz2;               // [W] <|m_z|^2>
};

/* StatSpinStruct: the 5 x Ns x Ns correlators (StatSpinStruct.cpp:16-135), `corr` summed over
 * the walkers; stream 5 x 2 Ns^2 per walker */
class StatSpinStruct : public Quantity {
  public:
    explicit StatSpinStruct(FileManager *fm = nullptr) : Quantity("StatSpinStruct", fm) {}
    uint32_t Mask() const override { return GPVMC_MEAS_STATSPIN; }
    void init(SlaterDeterminant &amp) override;
    void save(SlaterDeterminant &amp) override;
    std::vector<std::complex<double>> corr;  // [5][Ns][Ns], averaged over all measurements
};

/* w_m / mean(w) from ln w_m, without overflow */
std::vector<double> RelativeWeights(const std::vector<double> &logw);

/* OverlapTrack (OverlapTrack.cpp:6-29): the weights of one walk's measurements relative to
 * their mean, w_m / mean(w), per walker (stream n_meas x 1); also the mean of ln w */
class OverlapTrack : public Quantity {
  public:
    explicit OverlapTrack(FileManager *fm = nullptr) : Quantity("OverlapTrack", fm) {}
    uint32_t Mask() const override { return GPVMC_MEAS_OVERLAP; }
    uint32_t Traces() const override { return TraceLogW; }
    void TakeTraces(int W, int64_t n_meas, const double *logw, const double *mz) override;
    void init(SlaterDeterminant &amp) override;
    void save(SlaterDeterminant &amp) override;
    std::vector<double> mean_logw;         // [W]
    std::vector<std::vector<double>> rel;  // [W][n_meas] w_m / mean(w)

  private:
    std::vector<std::vector<double>> logw_;
};

/* StagMagnTrack (StagMagnTrack.cpp:17-50): m_z of every measurement of one walk, per walker
 * (stream 2 n_meas x 1); ground-state contexts */
class StagMagnTrack : public Quantity {
  public:
    explicit StagMagnTrack(FileManager *fm = nullptr) : Quantity("StagMagnTrack", fm) {}
    uint32_t Mask() const override { return GPVMC_MEAS_STAGMAGNZ; }
    uint32_t Traces() const override { return TraceStagMagn; }
    void TakeTraces(int W, int64_t n_meas, const double *logw, const double *mz) override;
    void init(SlaterDeterminant &amp) override;
    void save(SlaterDeterminant &amp) override;
    std::vector<std::vector<std::complex<double>>> mz;  // [W][n_meas]
};

class MetroMC {
  public:
    explicit MetroMC(LatticeStepper &stepper) : st_(stepper) {}
    void AddQuantity(Quantity *q) { quantities_.push_back(q); }
    const std::vector<Quantity *> &GetQuantities() const { return quantities_; }
    /* MetroMC::Walk(len, meas): len steps of every walker; with meas > 0 the quantities are
     * init()ed first and measured after step s when meas == 1 or s % meas == 0 */
    void Walk(int64_t len, int64_t meas);
    /* rejected / proposed steps of the measuring walks so far (MetroMC.cpp:97-109: steps and
     * rejections are counted only while measuring), over every walker, or per walker */
    double Rejection() const;
    std::vector<double> RejectionPerChain() const;

  private:
    LatticeStepper &st_;
    std::vector<Quantity *> quantities_;
    std::vector<double> steps_, rejected_;
};

}  // namespace gpvmc

#endif
