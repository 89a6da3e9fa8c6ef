// C++ host classes over the C ABI (include/gpvmc_host.hpp): the reference's object graph
// for a batch of chains on one GPU.  Host-side work only: the orbital table, the channel
// states and their hop graph, initial configurations, and the calls into libgpvmc_hip.so.
#include "gpvmc_host.hpp"

#include <hdf5.h>
#include <hdf5_hl.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <utility>

namespace gpvmc {

namespace {

void check(int rc)
{
    if (rc != GPVMC_OK) throw Error(gpvmc_last_error());
}

struct QN {
    int kx, ky, band;
};

// StagFluxWaveFunction::inmbz (StagFluxWaveFunction.cpp:89-101)
bool inmbz(int L, double bcx, double bcy, int kx, int ky)
{
    const double kkx = (kx + bcx / 2.0) / L, kky = (ky + bcy / 2.0) / L, eps = 1e-9;
    return (0 <= kkx && kkx < 1 && 0 <= kky && kky < 1) &&
           (kkx + kky - 0.5 < 0 || kkx + kky - 1.5 >= -eps || kky - kkx - 0.5 > eps || kky - kkx + 0.5 <= eps);
}

// orbital quantum numbers, kx outer and ky inner over the magnetic zone, two bands each
// (StagFluxWaveFunction.cpp:24-41)
std::vector<QN> magnetic_zone(int L, double bcx, double bcy)
{
    std::vector<QN> qn;
    for (int kx = 0; kx < L; ++kx)
        for (int ky = 0; ky < L; ++ky)
            if (inmbz(L, bcx, bcy, kx, ky)) {
                qn.push_back({kx, ky, 0});
                qn.push_back({kx, ky, 1});
            }
    return qn;
}

// StagFluxWaveFunction::mbzmod (StagFluxWaveFunction.cpp:103-120)
std::pair<int, int> mbzmod(int L, double bcx, double bcy, int kx, int ky)
{
    if (inmbz(L, bcx, bcy, kx, ky)) return {kx, ky};
    const int h = L / 2;
    if (kx < h && ky < h) return {kx + h, ky + h};
    if (kx < h && ky >= h) return {kx + h, ky - h};
    if (kx >= h && ky < h) return {kx - h, ky + h};
    return {kx - h, ky - h};
}

// Philox4x32-10 (Salmon et al., SC'11) with its Weyl key schedule
void philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

}  // namespace

StagFluxWaveFunction::StagFluxWaveFunction(int L, double phi, double neel, double neel_exp, double bcx, double bcy,
                                           Channel channel, int qx, int qy)
    : L_(L), Ns_(L * L), N_(L * L / 2), nk_(0), ch_(channel)
{
    if (L < 2 || (L * L) % 2) throw Error("L must be >= 2 with L*L even");
    const std::vector<QN> qn = magnetic_zone(L, bcx, bcy);
    if ((int)qn.size() != Ns_) throw Error("magnetic zone enumeration does not give L*L orbitals");
    for (const QN &q : qn) qn_.insert(qn_.end(), {q.kx, q.ky, q.band});
    // the orbital table (StagFluxWaveFunction.cpp:51-79, 138-163)
    const double pi = std::acos(-1.0), ph = phi * pi;
    table_.assign((size_t)2 * Ns_ * Ns_, 0.0);
    for (int f = 0; f < Ns_; ++f) {
        const double k0 = (qn[f].kx + bcx / 2.0) * 2 * pi / L, k1 = (qn[f].ky + bcy / 2.0) * 2 * pi / L;
        const std::complex<double> d =
            0.5 * (std::cos(k0) * std::polar(1.0, ph) + std::cos(k1) * std::polar(1.0, -ph));
        const double nk = neel_exp == 0.0 ? neel
                                          : neel * std::pow(0.5 * (std::cos(k0) * std::cos(k0) +
                                                                   std::cos(k1) * std::cos(k1)), neel_exp);
        const double ad = std::abs(d), om = std::sqrt(nk * nk + ad * ad);
        if (!(ad > 0.0)) throw Error("degenerate delta(k): choose bc or neel so that |delta| > 0");
        for (int fl = 0; fl < 2; ++fl) {
            const double sig = fl ? -1.0 : 1.0;
            const double up = std::sqrt(0.5 * (1 + sig * nk / om)), dn = std::sqrt(0.5 * (1 - sig * nk / om));
            const std::complex<double> U = qn[f].band == 0 ? std::complex<double>(up) : -std::conj(d) / ad * dn;
            const std::complex<double> V = qn[f].band == 0 ? d / ad * dn : std::complex<double>(up);
            for (int r = 0; r < Ns_; ++r) {
                const int ry = r / L, rx = r - ry * L;
                table_[((size_t)fl * Ns_ + f) * Ns_ + r] =
                    std::sqrt(2.0) * std::polar(1.0, k0 * rx + k1 * ry) * (((rx + ry) % 2 == 0) ? U : V);
            }
        }
    }
    // the channel's states, each flavour's occupied orbitals in increasing order
    std::vector<int> qn2fock((size_t)2 * Ns_, Ns_);
    for (int i = 0; i < Ns_; ++i) qn2fock[((size_t)qn[i].kx * L + qn[i].ky) * 2 + qn[i].band] = i;
    qx = ((qx % L) + L) % L;
    qy = ((qy % L) + L) % L;
    auto hole_of = [&](int fe) {
        const QN &o = qn[2 * fe + 1];
        const auto k = mbzmod(L, bcx, bcy, ((o.kx - qx) % L + L) % L, ((o.ky - qy) % L + L) % L);
        return qn2fock[((size_t)k.first * L + k.second) * 2];
    };
    std::vector<int> even(N_);
    for (int f = 0; f < N_; ++f) even[f] = 2 * f;
    std::vector<std::pair<std::vector<int>, std::vector<int>>> st;
    auto without = [&](int h) {
        std::vector<int> v;
        for (int o : even)
            if (o != h) v.push_back(o);
        return v;
    };
    if (channel == GroundState) {
        st.push_back({even, even});
    } else if (channel == TransExciton) {
        for (int fe = 0; fe < N_; ++fe) {
            std::vector<int> u = even;
            u.push_back(2 * fe + 1);
            st.push_back({u, without(hole_of(fe))});
        }
    } else {
        for (int fe = 0; fe < N_; ++fe) {
            std::vector<int> p = without(hole_of(fe));
            p.push_back(2 * fe + 1);
            st.push_back({p, even});
            st.push_back({even, p});
        }
        if ((qx == 0 && qy == 0) || (qx == L / 2 && qy == L / 2)) st.push_back({even, even});
    }
    nk_ = (int)st.size();
    npt_[0] = (int)st[0].first.size();
    npt_[1] = (int)st[0].second.size();
    for (auto &s : st) {
        std::sort(s.first.begin(), s.first.end());
        std::sort(s.second.begin(), s.second.end());
        std::vector<int32_t> v(s.first.begin(), s.first.end());
        v.insert(v.end(), s.second.begin(), s.second.end());
        states_.push_back(v);
        state_orb_.insert(state_orb_.end(), v.begin(), v.end());
    }
    // hop graph: every state is E = {0, 2, ..., 2N-2} with a few orbitals removed and added
    // per flavour; the path c -> k pairs the orbitals occupied in c but not in k with those
    // occupied in k but not in c, each in increasing order (State::get_hop_path,
    // State.cpp:213-231), and HopSign counts the occupied orbitals between the two ends of
    // every hop, hop after hop (State.cpp:86-115)
    std::vector<std::vector<int>> rem((size_t)nk_ * 2), add((size_t)nk_ * 2);
    for (int k = 0; k < nk_; ++k)
        for (int fl = 0; fl < 2; ++fl) {
            const std::vector<int> &o = fl ? st[k].second : st[k].first;
            for (int e : even)
                if (!std::binary_search(o.begin(), o.end(), e)) rem[k * 2 + fl].push_back(e);
            for (int x : o)
                if (x % 2 == 1 || x >= 2 * N_) add[k * 2 + fl].push_back(x);
        }
    auto has = [](const std::vector<int> &v, int x) { return std::find(v.begin(), v.end(), x) != v.end(); };
    auto n_even_between = [&](int lo, int hi) {
        const int a = lo + 1, b = std::min(hi - 1, 2 * N_ - 2);
        if (b < a) return 0;
        return b / 2 - (a + 1) / 2 + 1;
    };
    hop_n_.assign((size_t)nk_ * nk_ * 2, 0);
    hop_orb_.assign((size_t)nk_ * nk_ * 8, 0);
    hopsign_.assign((size_t)nk_ * nk_, 1);
    for (int c = 0; c < nk_; ++c)
        for (int k = 0; k < nk_; ++k) {
            bool odd = false;
            for (int fl = 0; fl < 2; ++fl) {
                const auto &rc = rem[c * 2 + fl], &ac = add[c * 2 + fl], &rk = rem[k * 2 + fl], &ak = add[k * 2 + fl];
                std::vector<int> holes, parts;
                for (int o : ac)
                    if (!has(ak, o)) holes.push_back(o);
                for (int o : rk)
                    if (!has(rc, o)) holes.push_back(o);
                for (int o : ak)
                    if (!has(ac, o)) parts.push_back(o);
                for (int o : rc)
                    if (!has(rk, o)) parts.push_back(o);
                std::sort(holes.begin(), holes.end());
                std::sort(parts.begin(), parts.end());
                if (holes.size() != parts.size() || holes.size() > 2)
                    throw Error("channel paths must have at most two hops per flavour");
                const size_t base = ((size_t)c * nk_ + k) * 2 + fl;
                hop_n_[base] = (int32_t)holes.size();
                std::vector<int> gone = rc, came = ac;
                for (size_t h = 0; h < holes.size(); ++h) {
                    const int a = holes[h], b = parts[h];
                    hop_orb_[(base * 2 + h) * 2] = a;
                    hop_orb_[(base * 2 + h) * 2 + 1] = b;
                    const int lo = std::min(a, b), hi = std::max(a, b);
                    int nb = n_even_between(lo, hi);
                    for (int o : gone)
                        if (lo < o && o < hi) --nb;
                    for (int o : came)
                        if (lo < o && o < hi) ++nb;
                    odd ^= (nb & 1) != 0;
                    auto ia = std::find(came.begin(), came.end(), a);
                    if (ia != came.end())
                        came.erase(ia);
                    else
                        gone.push_back(a);
                    auto ib = std::find(gone.begin(), gone.end(), b);
                    if (ib != gone.end())
                        gone.erase(ib);
                    else
                        came.push_back(b);
                }
            }
            hopsign_[(size_t)c * nk_ + k] = odd ? -1 : 1;
        }
}

std::vector<std::pair<int32_t, int32_t>> StagFluxWaveFunction::GetHop(int c, int k, int fl) const
{
    std::vector<std::pair<int32_t, int32_t>> p;
    const size_t base = ((size_t)c * nk_ + k) * 2 + fl;
    for (int h = 0; h < hop_n_[base]; ++h) p.push_back({hop_orb_[(base * 2 + h) * 2], hop_orb_[(base * 2 + h) * 2 + 1]});
    return p;
}

gpvmc_channel StagFluxWaveFunction::GraphABI() const
{
    gpvmc_channel ch;
    ch.n_up = npt_[0];
    ch.n_dn = npt_[1];
    ch.nk = nk_;
    ch.state_orb = state_orb_.data();
    ch.hop_n = hop_n_.data();
    ch.hop_orb = hop_orb_.data();
    ch.hopsign = hopsign_.data();
    return ch;
}

std::vector<int32_t> RanInitDraw(int L, int n_up, const std::function<double()> &uniform, int64_t *used)
{
    const int Ns = L * L;
    if (n_up < 0 || n_up > Ns) throw Error("RanInit: n_up out of range");
    std::vector<int32_t> s(Ns, -1);
    int64_t n_used = 0;
    for (int np = 0; np != n_up;) {
        const double u = uniform();
        ++n_used;
        const int v = (int)(u * Ns);
        if (v < 0 || v >= Ns) throw Error("RanInit: uniform outside [0, 1)");
        if (s[v] < 0) {
            s[v] = 0;
            ++np;
        }
    }
    for (int v = 0; v < Ns; ++v)
        if (s[v] < 0) s[v] = 1;
    if (used) *used = n_used;
    return s;
}

std::vector<int32_t> RanInit(int L, int n_walkers, int n_up, uint64_t seed, int attempt)
{
    const int Ns = L * L;
    std::vector<int32_t> out((size_t)n_walkers * Ns);
    for (int w = 0; w < n_walkers; ++w) {
        uint32_t i = 0;
        const std::vector<int32_t> s = RanInitDraw(L, n_up, [&]() {
            uint32_t c[4] = {i++, (uint32_t)attempt, (uint32_t)w, 2u};
            philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
            return ((double)c[0] + 0.5) * 0x1p-32;
        });
        std::copy(s.begin(), s.end(), out.begin() + (size_t)w * Ns);
    }
    return out;
}

std::vector<int32_t> RandomSpins(int L, int n_walkers, int n_up, uint64_t seed)
{
    const int Ns = L * L;
    std::vector<int32_t> out((size_t)n_walkers * Ns);
    for (int w = 0; w < n_walkers; ++w) {
        int32_t *s = out.data() + (size_t)w * Ns;
        for (int i = 0; i < Ns; ++i) s[i] = i < n_up ? 0 : 1;
        for (int i = Ns - 1; i > 0; --i) {
            uint32_t c[4] = {(uint32_t)i, 0u, (uint32_t)w, 1u};
            philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
            const int j = (int)((double)c[0] * 0x1p-32 * (i + 1));
            std::swap(s[i], s[j]);
        }
    }
    return out;
}

SlaterDeterminant::SlaterDeterminant(const StagFluxWaveFunction &wav, int n_walkers, int precision, uint64_t seed,
                                     int device)
    : wav_(wav), W_(n_walkers)
{
    const int L = wav.GetL(), Ns = L * L;
    gpvmc_config cfg = {L, n_walkers, 1, precision, device, 0, seed};
    const std::vector<std::complex<double>> &t = wav.Table();
    if (wav.GetChannel() == StagFluxWaveFunction::GroundState) {
        check(gpvmc_create(&cfg, &ctx_));
        // occ[fl][site][p] = Phi_fl(orbital of particle p = 2p, site) (StagFluxGroundState.cpp:11-14)
        const int N = Ns / 2;
        std::vector<double> occ((size_t)2 * Ns * N * 2);
        for (int fl = 0; fl < 2; ++fl)
            for (int r = 0; r < Ns; ++r)
                for (int p = 0; p < N; ++p) {
                    const std::complex<double> v = t[((size_t)fl * Ns + 2 * p) * Ns + r];
                    occ[(((size_t)fl * Ns + r) * N + p) * 2] = v.real();
                    occ[(((size_t)fl * Ns + r) * N + p) * 2 + 1] = v.imag();
                }
        try {
            check(gpvmc_set_orbitals(ctx_, 0, occ.data()));
        } catch (...) {
            gpvmc_destroy(ctx_);
            throw;
        }
    } else {
        const gpvmc_channel ch = wav.GraphABI();
        check(gpvmc_create_channel(&cfg, &ch, &ctx_));
        try {
            check(gpvmc_set_orbital_table(ctx_, 0, reinterpret_cast<const double *>(t.data())));
        } catch (...) {
            gpvmc_destroy(ctx_);
            throw;
        }
    }
}

SlaterDeterminant::~SlaterDeterminant()
{
    if (ctx_) gpvmc_destroy(ctx_);
}

std::vector<int32_t> SlaterDeterminant::Init(const std::vector<int32_t> &spins)
{
    const size_t Ns = (size_t)wav_.GetL() * wav_.GetL();
    if (spins.size() != (size_t)W_ * Ns) throw Error("spins must hold n_walkers * L * L values");
    std::vector<int32_t> ok(W_);
    check(gpvmc_init_walkers(ctx_, spins.data(), ok.data()));
    return ok;
}

std::vector<int32_t> SlaterDeterminant::InitRandom(uint64_t seed, int max_attempts, std::vector<int32_t> *attempts)
{
    std::vector<int32_t> ok(W_), att(W_);
    check(gpvmc_init_random(ctx_, seed, max_attempts, ok.data(), att.data()));
    if (attempts) *attempts = att;
    return ok;
}

void SlaterDeterminant::Refresh() { check(gpvmc_refresh(ctx_, nullptr)); }

std::vector<LogAmp> SlaterDeterminant::Amp() const
{
    std::vector<double> la(W_), ph(W_);
    check(gpvmc_read_state(ctx_, nullptr, nullptr, la.data(), ph.data()));
    std::vector<LogAmp> out(W_);
    for (int w = 0; w < W_; ++w) out[w] = {la[w], ph[w]};
    return out;
}

std::vector<int32_t> SlaterDeterminant::Spins() const
{
    std::vector<int32_t> s((size_t)W_ * wav_.GetL() * wav_.GetL());
    check(gpvmc_read_state(ctx_, s.data(), nullptr, nullptr, nullptr));
    return s;
}

void Quantity::init(SlaterDeterminant &amp) { check(gpvmc_reset_accumulators(amp.Context())); }

namespace {
std::vector<double> accumulators(SlaterDeterminant &amp)
{
    std::vector<double> acc((size_t)amp.NWalkers() * GPVMC_ACC_STRIDE);
    check(gpvmc_read_accumulators(amp.Context(), acc.data()));
    return acc;
}
}  // namespace

void ProjHeis::save(SlaterDeterminant &amp)
{
    const int W = amp.NWalkers();
    const std::vector<double> acc = accumulators(amp);
    if (amp.GetWaveFunction().GetChannel() == StagFluxWaveFunction::GroundState) {
        energy.assign(W, 0.0);
        for (int w = 0; w < W; ++w) {
            const double *a = acc.data() + (size_t)w * GPVMC_ACC_STRIDE;
            if (a[3] > 0) energy[w] = std::complex<double>(a[0], a[1]) / a[3];
            if (fm_) {
                // MatrixQuantity(2, 1): Val(0,0) = E_loc per site, Val(0,1) (= m_vals[1]) counts
                // the measurements, so its mean is 1 (ProjHeis.cpp:140-145)
                const double v[4] = {energy[w].real(), energy[w].imag(), a[3] > 0 ? 1.0 : 0.0, 0.0};
                fm_->FileStream(name_, w).Write(2, 2, v);
            }
        }
        return;
    }
    nk = amp.GetWaveFunction().GetNExc();
    const size_t kk = (size_t)nk * nk;
    std::vector<double> ho((size_t)W * 4 * kk);
    check(gpvmc_read_channel(amp.Context(), 0, W, ho.data(), nullptr, nullptr));
    H.assign((size_t)W * kk, 0.0);
    O.assign((size_t)W * kk, 0.0);
    std::vector<double> v(4 * kk);
    for (int w = 0; w < W; ++w) {
        const double n = acc[(size_t)w * GPVMC_ACC_STRIDE + 3];
        const double *b = ho.data() + (size_t)w * 4 * kk;
        if (n > 0)
            for (size_t e = 0; e < kk; ++e) {
                H[(size_t)w * kk + e] = std::complex<double>(b[2 * e], b[2 * e + 1]) / n;
                O[(size_t)w * kk + e] = std::complex<double>(b[2 * (kk + e)], b[2 * (kk + e) + 1]) / n;
            }
        if (fm_) {
            // rows k1 < nk: H[k1][k2]; rows nk + k1: O[k1][k2]; complex interleaved along rows
            for (size_t e = 0; e < kk; ++e) {
                v[2 * e] = H[(size_t)w * kk + e].real();
                v[2 * e + 1] = H[(size_t)w * kk + e].imag();
                v[2 * (kk + e)] = O[(size_t)w * kk + e].real();
                v[2 * (kk + e) + 1] = O[(size_t)w * kk + e].imag();
            }
            fm_->FileStream(name_, w).Write(2 * nk, 2 * nk, v.data());
        }
    }
}

void StagMagnZ::save(SlaterDeterminant &amp)
{
    const int W = amp.NWalkers();
    const std::vector<double> acc = accumulators(amp);
    mz.assign(W, 0.0);
    mz2.assign(W, 0.0);
    for (int w = 0; w < W; ++w) {
        const double *a = acc.data() + (size_t)w * GPVMC_ACC_STRIDE;
        if (a[3] > 0) {
            mz[w] = std::complex<double>(a[4], a[5]) / a[3];
            mz2[w] = a[6] / a[3];
        }
        if (fm_) {
            const double v[2] = {mz[w].real(), mz[w].imag()};
            fm_->FileStream(name_, w).Write(1, 2, v);
        }
    }
}

void StatSpinStruct::init(SlaterDeterminant &amp) { Quantity::init(amp); }

void StatSpinStruct::save(SlaterDeterminant &amp)
{
    const size_t Ns = (size_t)amp.GetWaveFunction().GetL() * amp.GetWaveFunction().GetL();
    const size_t blk = 5 * Ns * Ns * 2;
    std::vector<double> raw(blk);
    double n = 0;
    check(gpvmc_read_statspin_groups(amp.Context(), raw.data(), &n));
    corr.assign(5 * Ns * Ns, 0.0);
    if (n > 0)
        for (size_t e = 0; e < corr.size(); ++e) corr[e] = std::complex<double>(raw[2 * e], raw[2 * e + 1]) / n;
    if (!fm_) return;
    // per walker: MatrixQuantity(5, Ns^2) divided by its own count (StatSpinStruct.cpp:12-14),
    // read in chunks of walkers
    const int W = amp.NWalkers(), chunk = (int)std::max<size_t>(1, std::min<size_t>(256, (64u << 20) / (blk * 8)));
    std::vector<double> part((size_t)chunk * blk), cnt(chunk);
    for (int w0 = 0; w0 < W; w0 += chunk) {
        const int nw = std::min(chunk, W - w0);
        check(gpvmc_read_statspin(amp.Context(), w0, nw, part.data(), cnt.data()));
        for (int k = 0; k < nw; ++k) {
            double *b = part.data() + (size_t)k * blk;
            if (cnt[k] > 0)
                for (size_t e = 0; e < blk; ++e) b[e] /= cnt[k];
            fm_->FileStream(name_, w0 + k).Write(5, (int)(2 * Ns * Ns), b);
        }
    }
}

std::vector<double> RelativeWeights(const std::vector<double> &logw)
{
    // ln mean(w) = max + ln(sum_m exp(ln w_m - max) / n): no overflow where the reference
    // needs BigDouble (OverlapTrack.cpp:15-25)
    std::vector<double> rel(logw.size());
    if (logw.empty()) return rel;
    const double mx = *std::max_element(logw.begin(), logw.end());
    double sum = 0;
    for (double x : logw) sum += std::exp(x - mx);
    const double lmean = mx + std::log(sum / (double)logw.size());
    for (size_t m = 0; m < logw.size(); ++m) rel[m] = std::exp(logw[m] - lmean);
    return rel;
}

void OverlapTrack::init(SlaterDeterminant &amp)
{
    Quantity::init(amp);
    logw_.assign(amp.NWalkers(), {});
}

void OverlapTrack::TakeTraces(int W, int64_t n_meas, const double *logw, const double *)
{
    if ((int)logw_.size() != W) logw_.assign(W, {});
    for (int w = 0; w < W; ++w) logw_[w].insert(logw_[w].end(), logw + (size_t)w * n_meas, logw + (size_t)(w + 1) * n_meas);
}

void OverlapTrack::save(SlaterDeterminant &amp)
{
    const int W = amp.NWalkers();
    const std::vector<double> acc = accumulators(amp);
    mean_logw.assign(W, 0.0);
    rel.assign(W, {});
    for (int w = 0; w < W; ++w) {
        const double *a = acc.data() + (size_t)w * GPVMC_ACC_STRIDE;
        if (a[3] > 0) mean_logw[w] = a[11] / a[3];
        if (w < (int)logw_.size()) rel[w] = RelativeWeights(logw_[w]);
        if (fm_) fm_->FileStream(name_, w).Write((int)rel[w].size(), 1, rel[w].data());
    }
}

void StagMagnTrack::init(SlaterDeterminant &amp)
{
    Quantity::init(amp);
    mz.assign(amp.NWalkers(), {});
}

void StagMagnTrack::TakeTraces(int W, int64_t n_meas, const double *, const double *m)
{
    if ((int)mz.size() != W) mz.assign(W, {});
    for (int w = 0; w < W; ++w)
        for (int64_t k = 0; k < n_meas; ++k)
            mz[w].emplace_back(m[2 * ((size_t)w * n_meas + k)], m[2 * ((size_t)w * n_meas + k) + 1]);
}

void StagMagnTrack::save(SlaterDeterminant &amp)
{
    if (!fm_) return;
    for (int w = 0; w < amp.NWalkers(); ++w) {
        const std::vector<std::complex<double>> &v = w < (int)mz.size() ? mz[w] : std::vector<std::complex<double>>();
        fm_->FileStream(name_, w).Write((int)(2 * v.size()), 1, reinterpret_cast<const double *>(v.data()));
    }
}

void MetroMC::Walk(int64_t len, int64_t meas)
{
    SlaterDeterminant &amp = st_.GetAmp();
    uint32_t mask = 0, traces = 0;
    for (Quantity *q : quantities_) {
        mask |= q->Mask();
        traces |= q->Traces();
    }
    if (meas > 0 && mask == 0) meas = 0;
    if (meas > 0)
        for (Quantity *q : quantities_) q->init(amp);
    const int W = amp.NWalkers();
    const int64_t n_meas = meas > 0 ? (meas == 1 ? len : (len + meas - 1) / meas) : 0;
    std::vector<double> logw, mz;
    if (n_meas && (traces & TraceLogW)) logw.resize((size_t)W * n_meas);
    if (n_meas && (traces & TraceStagMagn)) mz.resize((size_t)W * n_meas * 2);
    const bool channel = amp.GetWaveFunction().GetChannel() != StagFluxWaveFunction::GroundState;
    if (channel && !mz.empty()) {
        // a channel context measures ProjHeis and OverlapTrack on the device but has no m_z
        // trace: walk each measured step on its own, then read the configuration it measured
        // (StagMagnTrack.cpp:17-50, m_z = (1/Ns) sum_v e^{i pi (x+y)} (2 up_v - 1)); the
        // Philox counters run on across the calls, so the chain is the one a single call walks
        const int L = amp.GetWaveFunction().GetL(), Ns = L * L;
        std::vector<double> lw1(W);
        int64_t s = 0, m = 0;
        while (s < len) {
            const bool measured = meas == 1 || s % meas == 0;
            const int64_t seg = measured ? 1 : std::min<int64_t>(len - s, meas - s % meas);
            check(gpvmc_walk(amp.Context(), seg, measured ? 1 : 0, mask, nullptr, nullptr, nullptr, nullptr, nullptr,
                             nullptr, measured && !logw.empty() ? lw1.data() : nullptr));
            if (measured) {
                const std::vector<int32_t> sp = amp.Spins();
                for (int w = 0; w < W; ++w) {
                    double mr = 0;
                    for (int v = 0; v < Ns; ++v) {
                        const int x = v % L, y = v / L;
                        mr += ((x + y) % 2 ? -1.0 : 1.0) * (sp[(size_t)w * Ns + v] == 0 ? 1.0 : -1.0);
                    }
                    mz[((size_t)w * n_meas + m) * 2] = mr / Ns;
                    mz[((size_t)w * n_meas + m) * 2 + 1] = 0.0;
                    if (!logw.empty()) logw[(size_t)w * n_meas + m] = lw1[w];
                }
                ++m;
            }
            s += seg;
        }
    } else {
        check(gpvmc_walk(amp.Context(), len, meas, mask, nullptr, nullptr, nullptr, nullptr, nullptr,
                         mz.empty() ? nullptr : mz.data(), logw.empty() ? nullptr : logw.data()));
    }
    if (n_meas && traces)
        for (Quantity *q : quantities_)
            if (q->Traces()) q->TakeTraces(W, n_meas, logw.empty() ? nullptr : logw.data(), mz.empty() ? nullptr : mz.data());
    if (meas > 0) {
        // the accumulators were cleared by init: they hold this walk's counts
        const std::vector<double> acc = accumulators(amp);
        steps_.resize(W, 0.0);
        rejected_.resize(W, 0.0);
        for (int w = 0; w < W; ++w) {
            steps_[w] += acc[(size_t)w * GPVMC_ACC_STRIDE + 7];
            rejected_[w] += acc[(size_t)w * GPVMC_ACC_STRIDE + 9];
        }
    }
}

double MetroMC::Rejection() const
{
    double rej = 0, steps = 0;
    for (size_t w = 0; w < steps_.size(); ++w) {
        rej += rejected_[w];
        steps += steps_[w];
    }
    return steps > 0 ? rej / steps : 0.0;
}

std::vector<double> MetroMC::RejectionPerChain() const
{
    std::vector<double> r(steps_.size(), 0.0);
    for (size_t w = 0; w < r.size(); ++w)
        if (steps_[w] > 0) r[w] = rejected_[w] / steps_[w];
    return r;
}

void MatStream::Write(int m, int n, const double *mat)
{
    if (m < 0 || n < 0) throw Error("MatStream::Write: negative size");
    const size_t sz = (size_t)m * n;
    if (stat_ == 0) {
        nrow_ = m;
        ncol_ = n;
        mat_.assign(mat, mat + sz);
        stat_ = 1;
        return;
    }
    if (m != nrow_ || n != ncol_) throw Error("MatStream::Write: sizes must match.");
    for (size_t a = 0; a < sz; ++a) mat_[a] = (stat_ * mat_[a] + mat[a]) / (stat_ + 1);
    ++stat_;
}

FileManager::FileManager(int n_chains) : n_chains_(n_chains)
{
    if (n_chains < 1) throw Error("FileManager: n_chains must be >= 1");
}

MatStream &FileManager::FileStream(const std::string &name, int chain)
{
    if (chain < 0 || chain >= n_chains_) throw Error("FileManager::FileStream: chain out of range");
    std::vector<MatStream> &v = streams_[name];
    if (v.empty()) v.resize(n_chains_);
    return v[chain];
}

void FileManager::DataAttribute(const std::string &key, double v) { dattr_[key].assign(n_chains_, v); }

void FileManager::DataAttribute(const std::string &key, const std::vector<double> &per_chain)
{
    if ((int)per_chain.size() != n_chains_) throw Error("FileManager::DataAttribute: one value per chain");
    dattr_[key] = per_chain;
}

void FileManager::Write()
{
    for (auto &it : streams_) {
        std::vector<std::vector<Dataset>> &rec = data_[it.first];
        if (rec.empty()) rec.resize(n_chains_);
        std::vector<int> written(n_chains_);
        for (int c = 0; c < n_chains_; ++c) {
            // a stream with nothing saved since its dataset closed (Stat() == 0) rewrites that
            // dataset, attributes included, as the reference does
            MatStream &ms = it.second[c];
            std::vector<Dataset> &ds = rec[c];
            if (ms.Stat() == 1 || ds.empty()) ds.emplace_back();
            Dataset &d = ds.back();
            d.rows = ms.Rows();
            d.cols = ms.Cols();
            d.mat = ms.Mat();
            for (const auto &a : dattr_) d.attr[a.first] = a.second[c];
            written[c] = (int)ds.size() - 1;
            if (ms.Stat() == stat_per_sample_) ms.Stat() = 0;
        }
        if (!prefix_.empty()) WriteHDF5(it.first, rec, written);
    }
}

std::vector<std::string> FileManager::Names() const
{
    std::vector<std::string> n;
    for (const auto &it : data_) n.push_back(it.first);
    return n;
}

const std::vector<FileManager::Dataset> &FileManager::Datasets(const std::string &name, int chain) const
{
    auto it = data_.find(name);
    if (it == data_.end() || chain < 0 || chain >= n_chains_) throw Error("FileManager::Datasets: no such record");
    return it->second[chain];
}

namespace {

// closes an HDF5 id when it leaves scope
struct H5Id {
    hid_t id;
    herr_t (*close)(hid_t);
    ~H5Id()
    {
        if (id >= 0) close(id);
    }
};

bool file_exists(const std::string &path)
{
    struct stat sb;
    return stat(path.c_str(), &sb) == 0;
}

void root_attributes(hid_t f, const std::map<std::string, double> &num, const std::map<std::string, std::string> &str,
                     const std::string &type)
{
    for (const auto &a : num)
        if (H5LTset_attribute_double(f, "/", a.first.c_str(), &a.second, 1) < 0)
            throw Error("HDF5: cannot set attribute " + a.first);
    for (const auto &a : str)
        if (H5LTset_attribute_string(f, "/", a.first.c_str(), a.second.c_str()) < 0)
            throw Error("HDF5: cannot set attribute " + a.first);
    if (H5LTset_attribute_string(f, "/", "type", type.c_str()) < 0) throw Error("HDF5: cannot set attribute type");
}

hid_t open_group(hid_t f, const std::string &g)
{
    if (H5Lexists(f, g.c_str(), H5P_DEFAULT) > 0) return H5Gopen2(f, g.c_str(), H5P_DEFAULT);
    return H5Gcreate2(f, g.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
}

}  // namespace

void FileManager::SetOutput(const std::string &dir, const std::string &prefix)
{
    dir_ = dir;
    prefix_ = prefix;
}

std::string FileManager::RankGroup(int c) const
{
    return "/rank-" + std::to_string(n_chains_ == 1 ? 0 : c + 1);
}

int64_t FileManager::WriteSimple(const std::string &name) const
{
    if (prefix_.empty()) throw Error("FileManager::WriteSimple: no output set");
    const std::string path = dir_ + "/" + prefix_ + "-" + name + ".h5";
    const hid_t f = H5Fcreate(path.c_str(), H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) throw Error("Could not create file \"" + path + "\".");
    root_attributes(f, fattr_num_, fattr_str_, name);
    return (int64_t)f;
}

void FileManager::WriteHDF5(const std::string &name, const std::vector<std::vector<Dataset>> &rec,
                            const std::vector<int> &written) const
{
    const std::string path = dir_ + "/" + prefix_ + "-" + name + ".h5";
    hid_t f;
    if (!file_exists(path)) {
        f = H5Fcreate(path.c_str(), H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
        if (f >= 0) {
            H5Id guard{f, H5Fclose};
            root_attributes(f, fattr_num_, fattr_str_, name);
            const std::time_t t = std::time(nullptr);
            if (H5LTset_attribute_string(f, "/", "date", std::asctime(std::localtime(&t))) < 0)
                throw Error("HDF5: cannot set attribute date");
            guard.id = -1;
        }
    } else {
        f = H5Fopen(path.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
    }
    if (f < 0) throw Error("Could not create nor open file \"" + path + "\".");
    H5Id file{f, H5Fclose};
    for (int c = 0; c < n_chains_; ++c) {
        const std::string g = RankGroup(c);
        H5Id grp{open_group(f, g), H5Gclose};
        if (grp.id < 0) throw Error("HDF5: cannot create group " + g + " in " + path);
        const int k = written[c];
        const Dataset &d = rec[c][k];
        const std::string dn = g + "/data-" + std::to_string(k);
        const hsize_t dims[2] = {(hsize_t)d.rows, (hsize_t)d.cols};
        herr_t e;
        if (H5Lexists(f, dn.c_str(), H5P_DEFAULT) <= 0) {
            e = H5LTmake_dataset_double(f, dn.c_str(), 2, dims, d.mat.data());
        } else {
            H5Id did{H5Dopen2(f, dn.c_str(), H5P_DEFAULT), H5Dclose};
            e = did.id < 0 ? -1 : H5Dwrite(did.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, d.mat.data());
        }
        if (e < 0) throw Error("Could not write dataset in file \"" + path + "\".");
        for (const auto &a : d.attr)
            if (H5LTset_attribute_double(f, dn.c_str(), a.first.c_str(), &a.second, 1) < 0)
                throw Error("HDF5: cannot set attribute " + a.first);
    }
}

void StagFluxWaveFunction::Save(const FileManager &fm) const
{
    H5Id f{(hid_t)fm.WriteSimple("WaveFunction"), H5Fclose};
    for (int fl = 0; fl < 2; ++fl) {
        // states_<fl>[e][s] = 1 where state e occupies orbital s of flavour fl
        std::vector<int> fock((size_t)nk_ * Ns_, 0);
        for (int e = 0; e < nk_; ++e) {
            const std::vector<int32_t> &o = states_[e];
            const int off = fl ? npt_[0] : 0;
            for (int i = 0; i < npt_[fl]; ++i) fock[(size_t)e * Ns_ + o[off + i]] = 1;
        }
        const std::string name = "states_" + std::to_string(fl);
        const hsize_t dims[2] = {(hsize_t)nk_, (hsize_t)Ns_};
        if (H5LTmake_dataset_int(f.id, name.c_str(), 2, dims, fock.data()) < 0)
            throw Error("HDF5: cannot write " + name);
        H5Id ds{H5Dopen2(f.id, name.c_str(), H5P_DEFAULT), H5Dclose};
        const hsize_t dim = (hsize_t)Ns_;
        H5Id sp{H5Screate_simple(1, &dim, nullptr), H5Sclose};
        const char *keys[3] = {"kx", "ky", "band"};
        for (int q = 0; q < 3; ++q) {
            std::vector<unsigned int> v(Ns_);
            for (int i = 0; i < Ns_; ++i) v[i] = (unsigned int)qn_[3 * i + q];
            H5Id at{H5Acreate2(ds.id, keys[q], H5T_STD_U32BE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose};
            if (at.id < 0 || H5Awrite(at.id, H5T_NATIVE_UINT, v.data()) < 0)
                throw Error(std::string("HDF5: cannot write attribute ") + keys[q]);
        }
    }
}

void SaveFockState(const FileManager &fm, int L, const std::vector<int32_t> &spins)
{
    const int Ns = L * L, W = fm.NChains();
    if ((int64_t)spins.size() != (int64_t)W * Ns) throw Error("SaveFockState: spins must hold every chain");
    H5Id f{(hid_t)fm.WriteSimple("FockState"), H5Fclose};
    for (int fl = 0; fl < 2; ++fl)
        for (int c = 0; c < W; ++c) {
            std::vector<int> occ(Ns);
            for (int v = 0; v < Ns; ++v) occ[v] = spins[(size_t)c * Ns + v] == fl ? 1 : 0;
            H5Id grp{open_group(f.id, fm.RankGroup(c)), H5Gclose};
            const hsize_t dim = (hsize_t)Ns;
            if (grp.id < 0 || H5LTmake_dataset_int(grp.id, ("flavour-" + std::to_string(fl)).c_str(), 1, &dim,
                                                    occ.data()) < 0)
                throw Error("HDF5: cannot write the FockState of chain " + std::to_string(c));
        }
}

std::vector<int32_t> LoadFockState(const std::string &path, int L, int n_chains)
{
    const int Ns = L * L;
    H5Id f{H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose};
    if (f.id < 0) throw Error("cannot open spin state file " + path);
    FileManager names(n_chains);
    std::vector<int32_t> spins((size_t)n_chains * Ns, -1);
    for (int c = 0; c < n_chains; ++c)
        for (int fl = 0; fl < 2; ++fl) {
            const std::string dn = names.RankGroup(c) + "/flavour-" + std::to_string(fl);
            int rank = 0;
            hsize_t dim = 0;
            if (H5Lexists(f.id, names.RankGroup(c).c_str(), H5P_DEFAULT) <= 0 ||
                H5Lexists(f.id, dn.c_str(), H5P_DEFAULT) <= 0 || H5LTget_dataset_ndims(f.id, dn.c_str(), &rank) < 0 ||
                rank != 1 || H5LTget_dataset_info(f.id, dn.c_str(), &dim, nullptr, nullptr) < 0 || (int)dim != Ns)
                throw Error("spin state file " + path + " has no " + std::to_string(Ns) + "-site dataset " + dn);
            std::vector<int> occ(Ns);
            if (H5LTread_dataset_int(f.id, dn.c_str(), occ.data()) < 0) throw Error("cannot read " + dn);
            for (int v = 0; v < Ns; ++v)
                if (occ[v]) {
                    int32_t &s = spins[(size_t)c * Ns + v];
                    if (s >= 0) throw Error("spin state file " + path + ": site occupied twice in " + dn);
                    s = fl;
                }
        }
    for (int32_t s : spins)
        if (s < 0) throw Error("spin state file " + path + ": a site is empty");
    return spins;
}

}  // namespace gpvmc

// Host-side tables for the CPU tests (no device needed): the orbital table and the
// channel graph exactly as the C++ classes build them.
extern "C" int gpvmc_host_orbital_table(int32_t L, double phi, double neel, double neel_exp, double bcx, double bcy,
                                        double *out)
{
    try {
        gpvmc::StagFluxWaveFunction wf(L, phi, neel, neel_exp, bcx, bcy);
        const auto &t = wf.Table();
        for (size_t i = 0; i < t.size(); ++i) {
            out[2 * i] = t[i].real();
            out[2 * i + 1] = t[i].imag();
        }
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

// channel 0 ground state, 1 trans, 2 long; sizes first (out == NULL), then the arrays
extern "C" int gpvmc_host_channel(int32_t L, int32_t channel, int32_t qx, int32_t qy, int32_t *nk, int32_t *n_up,
                                  int32_t *n_dn, int32_t *state_orb, int32_t *hop_n, int32_t *hop_orb,
                                  int32_t *hopsign)
{
    try {
        gpvmc::StagFluxWaveFunction wf(L, 0.25, 0.1, 0.0, 1.0, 1.0, (gpvmc::StagFluxWaveFunction::Channel)channel, qx,
                                       qy);
        const gpvmc_channel ch = wf.GraphABI();
        *nk = ch.nk;
        *n_up = ch.n_up;
        *n_dn = ch.n_dn;
        const size_t K = ch.nk;
        if (state_orb) std::copy(ch.state_orb, ch.state_orb + K * (ch.n_up + ch.n_dn), state_orb);
        if (hop_n) std::copy(ch.hop_n, ch.hop_n + K * K * 2, hop_n);
        if (hop_orb) std::copy(ch.hop_orb, ch.hop_orb + K * K * 8, hop_orb);
        if (hopsign) std::copy(ch.hopsign, ch.hopsign + K * K, hopsign);
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

// Fisher-Yates configurations (the Python host's model.random_spins)
extern "C" int gpvmc_host_random_spins(int32_t L, int32_t n_walkers, int32_t n_up, uint64_t seed, int32_t *spins)
{
    const std::vector<int32_t> s = gpvmc::RandomSpins(L, n_walkers, n_up, seed);
    std::copy(s.begin(), s.end(), spins);
    return 0;
}

// RanInit's configurations from Philox (vmc_host's draw for attempt `attempt`)
extern "C" int gpvmc_host_raninit(int32_t L, int32_t n_walkers, int32_t n_up, uint64_t seed, int32_t attempt,
                                  int32_t *spins)
{
    const std::vector<int32_t> s = gpvmc::RanInit(L, n_walkers, n_up, seed, attempt);
    std::copy(s.begin(), s.end(), spins);
    return 0;
}

// RanInit's draw on a given uniform stream u[0 .. nu): returns the uniforms consumed, -1
// when the stream runs out
extern "C" int64_t gpvmc_host_raninit_stream(int32_t L, int32_t n_up, const double *u, int64_t nu, int32_t *spins)
{
    try {
        int64_t i = 0, used = 0;
        const std::vector<int32_t> s = gpvmc::RanInitDraw(L, n_up, [&]() {
            if (i >= nu) throw gpvmc::Error("out of uniforms");
            return u[i++];
        }, &used);
        std::copy(s.begin(), s.end(), spins);
        return used;
    } catch (const std::exception &) {
        return -1;
    }
}

// OverlapTrack's w_m / mean(w) from ln w_m
extern "C" int gpvmc_host_relative_weights(int32_t n, const double *logw, double *rel)
{
    const std::vector<double> r = gpvmc::RelativeWeights(std::vector<double>(logw, logw + n));
    std::copy(r.begin(), r.end(), rel);
    return 0;
}

// A FileManager driven through a scripted sequence of saves: save s writes mats[s][c]
// (m x n) into chain c's stream "Q", with data attribute "statistics" = s; Write() after
// every save.  Returns the number of datasets of each chain and their matrices
// out[c][k][m * n] (k < max_ds) and statistics attributes stats[c][k]; -1 on an error
// (a size mismatch when shapes[s] differs, message in gpvmc_host_last_error).
static std::string g_host_error;
extern "C" const char *gpvmc_host_last_error(void) { return g_host_error.c_str(); }
extern "C" int gpvmc_host_filemanager_script(int32_t n_chains, int32_t stat_per_sample, int32_t n_saves,
                                             const int32_t *shapes, const double *mats, int32_t max_ds,
                                             int32_t *n_ds, double *out, double *stats)
{
    try {
        gpvmc::FileManager fm(n_chains);
        fm.StatPerSample() = stat_per_sample;
        size_t off = 0, mn0 = (size_t)shapes[0] * shapes[1];
        for (int s = 0; s < n_saves; ++s) {
            const int m = shapes[2 * s], n = shapes[2 * s + 1];
            for (int c = 0; c < n_chains; ++c) {
                fm.FileStream("Q", c).Write(m, n, mats + off);
                off += (size_t)m * n;
            }
            fm.DataAttribute("statistics", (double)s);
            fm.Write();
        }
        for (int c = 0; c < n_chains; ++c) {
            const auto &ds = fm.Datasets("Q", c);
            n_ds[c] = (int32_t)ds.size();
            for (size_t k = 0; k < ds.size() && (int)k < max_ds; ++k) {
                std::copy(ds[k].mat.begin(), ds[k].mat.end(), out + ((size_t)c * max_ds + k) * mn0);
                stats[(size_t)c * max_ds + k] = ds[k].attr.at("statistics");
            }
        }
        return 0;
    } catch (const std::exception &e) {
        g_host_error = e.what();
        return -1;
    }
}

// The FileManager / WaveFunction::Save / SaveFockState script of the reference harness's
// h5layout mode, through these classes: prefix 7 (ground state at L, phi, neel; a
// ProjHeis-named stream written 4 times, 2 saves per sample; the FockState of `spins`)
// and prefix 8 (the transverse exciton's WaveFunction file at q = (1, 2)).
extern "C" int gpvmc_host_h5_script(const char *dir, int32_t L, double phi, double neel, const int32_t *spins)
{
    try {
        {
            gpvmc::FileManager fm(1);
            fm.SetOutput(dir, "7");
            fm.StatPerSample() = 2;
            fm.FileAttribute("L", (double)L);
            fm.FileAttribute("phi", phi);
            fm.FileAttribute("channel", std::string("groundstate"));
            gpvmc::StagFluxWaveFunction wav(L, phi, neel);
            wav.Save(fm);
            for (int s = 0; s < 4; ++s) {
                const double m[4] = {0.25 * s - 1.0, 1.0 / (s + 3.0), -0.5 * s, 2.0 + s};
                fm.FileStream("ProjHeis", 0).Write(2, 2, m);
                fm.DataAttribute("rej", 0.1 * s);
                fm.DataAttribute("time", 0.0);
                fm.DataAttribute("statistics", (double)((s % 2 + 1) * 100));
                fm.Write();
            }
            gpvmc::SaveFockState(fm, L, std::vector<int32_t>(spins, spins + L * L));
        }
        {
            gpvmc::FileManager fm(1);
            fm.SetOutput(dir, "8");
            fm.FileAttribute("L", (double)L);
            fm.FileAttribute("channel", std::string("trans"));
            gpvmc::StagFluxWaveFunction wav(L, phi, neel, 0.0, 1.0, 1.0, gpvmc::StagFluxWaveFunction::TransExciton, 1, 2);
            wav.Save(fm);
        }
        return 0;
    } catch (const std::exception &e) {
        g_host_error = e.what();
        return -1;
    }
}

// A double dataset of an HDF5 file: writes at most `max` values, returns how many it has
// (-1 when it cannot be read)
extern "C" int64_t gpvmc_host_h5_read(const char *path, const char *dataset, double *out, int64_t max)
{
    const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    int64_t n = -1;
    int rank = 0;
    if (H5Lexists(f, dataset, H5P_DEFAULT) > 0 && H5LTget_dataset_ndims(f, dataset, &rank) >= 0 && rank <= 4) {
        hsize_t dims[4] = {1, 1, 1, 1};
        if (H5LTget_dataset_info(f, dataset, dims, nullptr, nullptr) >= 0) {
            n = 1;
            for (int i = 0; i < rank; ++i) n *= (int64_t)dims[i];
            if (n <= max && H5LTread_dataset_double(f, dataset, out) < 0) n = -1;
        }
    }
    H5Fclose(f);
    return n;
}

// A scalar double attribute of an object of an HDF5 file (0 on success)
extern "C" int gpvmc_host_h5_attr(const char *path, const char *object, const char *attr, double *out)
{
    const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    const int rc = H5LTget_attribute_double(f, object, attr, out) < 0 ? -1 : 0;
    H5Fclose(f);
    return rc;
}

// FockState files for tests: write spins[w][site] of n_chains chains to
// <dir>/<prefix>-FockState.h5, or read them back (the --spinstate reader)
extern "C" int gpvmc_host_fockstate_write(const char *dir, const char *prefix, int32_t L, int32_t n_chains,
                                          const int32_t *spins)
{
    try {
        gpvmc::FileManager fm(n_chains);
        fm.SetOutput(dir, prefix);
        gpvmc::SaveFockState(fm, L, std::vector<int32_t>(spins, spins + (size_t)n_chains * L * L));
        return 0;
    } catch (const std::exception &e) {
        g_host_error = e.what();
        return -1;
    }
}

extern "C" int gpvmc_host_fockstate_read(const char *path, int32_t L, int32_t n_chains, int32_t *spins)
{
    try {
        const std::vector<int32_t> s = gpvmc::LoadFockState(path, L, n_chains);
        std::copy(s.begin(), s.end(), spins);
        return 0;
    } catch (const std::exception &e) {
        g_host_error = e.what();
        return -1;
    }
}

// The sample grouping of the merge: sample_of[i] and rank_in_sample[i] of record i (its
// sample and its place among that sample's members), totals[s] of each sample
extern "C" int gpvmc_host_bunch_records(int32_t n, const int32_t *statistics, int32_t nsamp, int32_t *sample_of,
                                        int32_t *rank_in_sample, int32_t *totals)
{
    try {
        const gpvmc::RecordBunches b = gpvmc::BunchRecords(nsamp, std::vector<int>(statistics, statistics + n));
        for (size_t s = 0; s < b.members.size(); ++s) {
            totals[s] = b.statistics[s];
            for (size_t i = 0; i < b.members[s].size(); ++i) {
                sample_of[b.members[s][i]] = (int32_t)s;
                rank_in_sample[b.members[s][i]] = (int32_t)i;
            }
        }
        return 0;
    } catch (const std::exception &e) {
        g_host_error = e.what();
        return -1;
    }
}

// gpvmc::Concatenate over n_in input paths
extern "C" int gpvmc_host_concatenate(int32_t n_in, const char *const *in, const char *out, int32_t nsamp)
{
    try {
        gpvmc::Concatenate(std::vector<std::string>(in, in + n_in), out, nsamp);
        return 0;
    } catch (const std::exception &e) {
        g_host_error = e.what();
        return -1;
    }
}

// Root attributes "key=value;key=value" (a value that reads whole as a number is numeric,
// otherwise a string) on a FileManager
static void set_root_attributes(gpvmc::FileManager &fm, const char *spec)
{
    std::string all = spec ? spec : "";
    size_t pos = 0;
    while (pos < all.size()) {
        size_t end = all.find(';', pos);
        if (end == std::string::npos) end = all.size();
        const std::string item = all.substr(pos, end - pos);
        const size_t eq = item.find('=');
        if (eq == std::string::npos || eq == 0) throw gpvmc::Error("bad root attribute \"" + item + "\"");
        const std::string key = item.substr(0, eq), val = item.substr(eq + 1);
        char *stop = nullptr;
        const double x = std::strtod(val.c_str(), &stop);
        if (!val.empty() && stop && *stop == 0)
            fm.FileAttribute(key, x);
        else
            fm.FileAttribute(key, val);
        pos = end + 1;
    }
}

// Records for tests of the merge and of the post-processing: a FileManager of n_chains
// chains writing `name` into <dir>/<prefix>-<name>.h5, one dataset per write (rows x cols
// doubles per chain, data [write][chain][rows*cols]) with the data attribute
// statistics[write], and the root attributes of root_attrs
extern "C" int gpvmc_host_records_write(const char *dir, const char *prefix, const char *name, const char *root_attrs,
                                        int32_t n_chains, int32_t n_writes, int32_t rows, int32_t cols,
                                        const double *data, const double *statistics)
{
    try {
        gpvmc::FileManager fm(n_chains);
        fm.SetOutput(dir, prefix);
        set_root_attributes(fm, root_attrs);
        const size_t mn = (size_t)rows * cols;
        for (int k = 0; k < n_writes; ++k) {
            for (int c = 0; c < n_chains; ++c)
                fm.FileStream(name, c).Write(rows, cols, data + ((size_t)k * n_chains + c) * mn);
            fm.DataAttribute("statistics", statistics[k]);
            fm.Write();
        }
        return 0;
    } catch (const std::exception &e) {
        g_host_error = e.what();
        return -1;
    }
}

// WaveFunction::Save of a channel into <dir>/<prefix>-WaveFunction.h5 with the root
// attributes of root_attrs (channel: 0 ground state, 1 trans, 2 long; phi in units of pi)
extern "C" int gpvmc_host_wavefunction_save(const char *dir, const char *prefix, const char *root_attrs, int32_t L,
                                            double phi, double neel, int32_t channel, int32_t qx, int32_t qy)
{
    try {
        gpvmc::FileManager fm(1);
        fm.SetOutput(dir, prefix);
        set_root_attributes(fm, root_attrs);
        gpvmc::StagFluxWaveFunction wav(L, phi, neel, 0.0, 1.0, 1.0, (gpvmc::StagFluxWaveFunction::Channel)channel,
                                        qx, qy);
        wav.Save(fm);
        return 0;
    } catch (const std::exception &e) {
        g_host_error = e.what();
        return -1;
    }
}

// ---- HDF5 inspection for host-side post-processing (the image has no h5py) ----------------
namespace {

herr_t collect_link(hid_t, const char *name, const H5L_info_t *, void *op)
{
    static_cast<std::vector<std::string> *>(op)->push_back(name);
    return 0;
}

herr_t collect_attr(hid_t, const char *name, const H5A_info_t *, void *op)
{
    static_cast<std::vector<std::string> *>(op)->push_back(name);
    return 0;
}

// names joined by '\n' into buf (cap bytes with the terminating zero); returns the bytes the
// whole list needs, so a caller can retry with a larger buffer
int64_t put_names(const std::vector<std::string> &names, char *buf, int64_t cap)
{
    std::string all;
    for (size_t i = 0; i < names.size(); ++i) all += (i ? "\n" : "") + names[i];
    if (buf && cap > 0) {
        const size_t n = std::min<size_t>(all.size(), (size_t)cap - 1);
        std::copy(all.begin(), all.begin() + n, buf);
        buf[n] = 0;
    }
    return (int64_t)all.size() + 1;
}

}  // namespace

// The members of a group (in name order): returns the buffer size the list needs, -1 when
// the file or the group cannot be opened
extern "C" int64_t gpvmc_host_h5_children(const char *path, const char *group, char *buf, int64_t cap)
{
    const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    std::vector<std::string> names;
    const herr_t rc = H5Literate_by_name(f, group, H5_INDEX_NAME, H5_ITER_INC, nullptr, collect_link, &names,
                                          H5P_DEFAULT);
    H5Fclose(f);
    return rc < 0 ? -1 : put_names(names, buf, cap);
}

// The shape of a dataset: its rank (dims[0 .. rank) written, at most max_rank), -1 on error
extern "C" int32_t gpvmc_host_h5_shape(const char *path, const char *dataset, int64_t *dims, int32_t max_rank)
{
    const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    int rank = -1;
    if (H5Lexists(f, dataset, H5P_DEFAULT) > 0 && H5LTget_dataset_ndims(f, dataset, &rank) >= 0 && rank <= max_rank) {
        std::vector<hsize_t> d(std::max(rank, 1));
        if (H5LTget_dataset_info(f, dataset, d.data(), nullptr, nullptr) < 0)
            rank = -1;
        else
            for (int i = 0; i < rank; ++i) dims[i] = (int64_t)d[i];
    } else {
        rank = -1;
    }
    H5Fclose(f);
    return rank;
}

// The attribute names of an object ("/" for the file), as gpvmc_host_h5_children
extern "C" int64_t gpvmc_host_h5_attr_names(const char *path, const char *object, char *buf, int64_t cap)
{
    const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    std::vector<std::string> names;
    const herr_t rc = H5Aiterate_by_name(f, object, H5_INDEX_NAME, H5_ITER_INC, nullptr, collect_attr, &names,
                                         H5P_DEFAULT);
    H5Fclose(f);
    return rc < 0 ? -1 : put_names(names, buf, cap);
}

// An attribute's value: a string attribute is copied into str (strcap bytes with the zero)
// and -2 is returned; a numeric one is converted to doubles (at most maxn into num) and its
// element count is returned; -1 on error
extern "C" int64_t gpvmc_host_h5_attr_value(const char *path, const char *object, const char *name, double *num,
                                            int64_t maxn, char *str, int64_t strcap)
{
    const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    int64_t out = -1;
    const hid_t a = H5Aopen_by_name(f, object, name, H5P_DEFAULT, H5P_DEFAULT);
    if (a >= 0) {
        const hid_t type = H5Aget_type(a), space = H5Aget_space(a);
        const hssize_t n = H5Sget_simple_extent_npoints(space);
        if (H5Tget_class(type) == H5T_STRING) {
            const size_t len = H5Tget_size(type);
            if (!H5Tis_variable_str(type) && n == 1) {
                std::vector<char> s(len + 1, 0);
                if (H5Aread(a, type, s.data()) >= 0) {
                    const size_t k = std::min<size_t>(std::strlen(s.data()), strcap > 0 ? (size_t)strcap - 1 : 0);
                    if (str && strcap > 0) {
                        std::copy(s.data(), s.data() + k, str);
                        str[k] = 0;
                    }
                    out = -2;
                }
            }
        } else if (n >= 0) {
            std::vector<double> v((size_t)std::max<hssize_t>(n, 1));
            if (H5Aread(a, H5T_NATIVE_DOUBLE, v.data()) >= 0) {
                std::copy(v.begin(), v.begin() + std::min<int64_t>(n, maxn), num);
                out = n;
            }
        }
        H5Sclose(space);
        H5Tclose(type);
        H5Aclose(a);
    }
    H5Fclose(f);
    return out;
}
