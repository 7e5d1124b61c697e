// HIP backend of the drivers (see hip_backend.hpp).
#include "hip_backend.hpp"

#include <algorithm>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <cstring>
#include <ctime>
#include <numeric>
#include <random>
#include <sstream>
#include <stdexcept>

#include "logger.hpp"

BackendOptions &BackendOptions::get() {
    static BackendOptions options;
    return options;
}

static hipStream_t stream_of(const std::shared_ptr<DeviceContext> &ctx) {
    auto *h = dynamic_cast<HipDeviceContext *>(ctx.get());
    if (!h) throw std::logic_error("the HIP backend needs a HipDeviceContext");
    return h->getStream();
}

// ---------------------------------------------------------------------------------
// Sample storage: F16 keeps samples/scale with a power-of-two scale per trace (FP16's
// range), so the scale itself adds no rounding; the kernels read float(stored) * scale.
// BF16 has FP32's exponent range, where such a scale changes no rounded bit: it keeps the
// samples themselves, with no scale.
static float pow2_scale(const float *x, unsigned n) {
    float amax = 0.0f;
    for (unsigned i = 0; i < n; i++) amax = std::max(amax, std::fabs(x[i]));
    if (!(amax > 0.0f) || !std::isfinite(amax)) return 1.0f;
    int e = 0;
    const float m = std::frexp(amax, &e);  // amax = m * 2^e, m in [0.5, 1)
    return std::ldexp(1.0f, m == 0.5f ? e - 1 : e);
}

static uint16_t to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return static_cast<uint16_t>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

static uint16_t to_f16(float f) {
    _Float16 h = static_cast<_Float16>(f);
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}

void HipLine::upload(const Gather &gather, int storage, hipStream_t stream) {
    const unsigned ntr = gather.getTotalTracesCount();
    const unsigned ns = gather.getSamplesPerTrace();
    mid.resize(ntr);
    hoff.resize(ntr);
    std::vector<float> hoff2(ntr);
    for (unsigned i = 0; i < ntr; i++) {
        mid[i] = gather.getMidpointOfTrace(i);
        hoff[i] = gather.getHalfoffsetOfTrace(i);
        hoff2[i] = hoff[i] * hoff[i];
    }
    const size_t n = static_cast<size_t>(ntr) * ns;
    if (storage == HSS_F32) {
        std::vector<float> s(n);
        for (unsigned i = 0; i < ntr; i++) {
            const auto &x = gather.getTraceAtIndex(i).getSamples();
            std::copy(x.begin(), x.end(), s.begin() + static_cast<std::ptrdiff_t>(i) * ns);
        }
        dSamples.resize(n * sizeof(float));
        dSamples.upload(s.data(), n * sizeof(float), stream);
        dScale.resize(0);
    } else {
        const bool scaled = storage == HSS_F16;
        std::vector<uint16_t> s(n);
        std::vector<float> scale(scaled ? ntr : 0);
        for (unsigned i = 0; i < ntr; i++) {
            const auto &x = gather.getTraceAtIndex(i).getSamples();
            const float sc = scaled ? pow2_scale(x.data(), ns) : 1.0f;
            if (scaled) scale[i] = sc;
            for (unsigned k = 0; k < ns; k++) {
                const float v = scaled ? x[k] / sc : x[k];
                s[static_cast<size_t>(i) * ns + k] = scaled ? to_f16(v) : to_bf16(v);
            }
        }
        dSamples.resize(n * sizeof(uint16_t));
        dSamples.upload(s.data(), n * sizeof(uint16_t), stream);
        dScale.resize(scale.size() * sizeof(float));
        if (scaled) dScale.upload(scale.data(), ntr * sizeof(float), stream);
    }
    dMid.resize(ntr * sizeof(float));
    dMid.upload(mid.data(), ntr * sizeof(float), stream);
    dHoff.resize(ntr * sizeof(float));
    dHoff.upload(hoff.data(), ntr * sizeof(float), stream);
    dHoff2.resize(ntr * sizeof(float));
    dHoff2.upload(hoff2.data(), ntr * sizeof(float), stream);
    view = hss_traces{};
    view.samples = dSamples.get();
    view.scale = storage == HSS_F16 ? dScale.as<float>() : nullptr;
    view.mid = dMid.as<float>();
    view.hoff = dHoff.as<float>();
    view.hoff2 = dHoff2.as<float>();
    view.ntr = static_cast<int32_t>(ntr);
    view.ns = static_cast<int32_t>(ns);
    view.dt = gather.getSamplePeriodInSeconds();
    view.tau = gather.getTauIndexDisplacement();
    view.w = static_cast<int32_t>(gather.getWindowSize());
    view.storage = storage;
    LOGD("Line resident on the device: " << ntr << " traces x " << ns << " samples");
}

// ---------------------------------------------------------------------------------
void MidpointBatch::clear() {
    m0.clear();
    beg.clear();
    cnt.clear();
    index.clear();
    indexed = false;
}

void MidpointBatch::addRange(float m, int begin, int count) {
    if (indexed) {
        // keep one addressing mode per batch: spell the range out
        m0.push_back(m);
        beg.push_back(static_cast<int32_t>(index.size()));
        cnt.push_back(count);
        for (int i = 0; i < count; i++) index.push_back(begin + i);
        return;
    }
    m0.push_back(m);
    beg.push_back(begin);
    cnt.push_back(count);
}

void MidpointBatch::addIndexList(float m, const std::vector<int32_t> &traces) {
    if (!indexed) {
        // convert the ranges added so far
        std::vector<int32_t> idx;
        for (size_t g = 0; g < beg.size(); g++) {
            const int32_t b = beg[g];
            beg[g] = static_cast<int32_t>(idx.size());
            for (int32_t i = 0; i < cnt[g]; i++) idx.push_back(b + i);
        }
        index = std::move(idx);
        indexed = true;
    }
    m0.push_back(m);
    beg.push_back(static_cast<int32_t>(index.size()));
    cnt.push_back(static_cast<int32_t>(traces.size()));
    index.insert(index.end(), traces.begin(), traces.end());
}

void MidpointBatch::upload(hipStream_t stream) {
    const size_t n = m0.size();
    dBeg.resize(std::max<size_t>(n, 1) * 4);
    dCnt.resize(std::max<size_t>(n, 1) * 4);
    dM0.resize(std::max<size_t>(n, 1) * 4);
    dBeg.upload(beg.data(), n * 4, stream);
    dCnt.upload(cnt.data(), n * 4, stream);
    dM0.upload(m0.data(), n * 4, stream);
    // every launch over the batch takes the midpoints up in the XCD-balanced order; the
    // results stay in batch order
    ord = xcd_balanced_order(cnt);
    dOrder.resize(std::max<size_t>(n, 1) * 4);
    dOrder.upload(ord.data(), n * 4, stream);
    if (indexed) {
        dIndex.resize(std::max<size_t>(index.size(), 1) * 4);
        dIndex.upload(index.data(), index.size() * 4, stream);
    }
}

void MidpointBatch::retain(const std::vector<size_t> &positions) {
    std::vector<float> m;
    std::vector<int32_t> b, c, idx;
    for (size_t p : positions) {
        m.push_back(m0.at(p));
        c.push_back(cnt.at(p));
        if (indexed) {
            b.push_back(static_cast<int32_t>(idx.size()));
            idx.insert(idx.end(), index.begin() + beg[p], index.begin() + beg[p] + cnt[p]);
        } else {
            b.push_back(beg[p]);
        }
    }
    m0 = std::move(m);
    beg = std::move(b);
    cnt = std::move(c);
    index = std::move(idx);
}

hss_midpoints MidpointBatch::view(float h0, float apm) const {
    hss_midpoints v{};
    v.n = static_cast<int32_t>(m0.size());
    v.begin = dBeg.as<int32_t>();
    v.count = dCnt.as<int32_t>();
    v.m0 = dM0.as<float>();
    v.index = indexed ? dIndex.as<int32_t>() : nullptr;
    v.h0 = h0;
    v.apm = apm;
    v.order = m0.size() > 1 ? dOrder.as<int32_t>() : nullptr;
    return v;
}

TraceRange selectRange(const std::vector<float> &m, traveltime_t model, float m0, float apm) {
    if (model == CMP) {
        // m == m0 (cmp/common.cu:13); the line is sorted by midpoint
        auto lo = std::lower_bound(m.begin(), m.end(), m0);
        auto hi = std::upper_bound(lo, m.end(), m0);
        return {static_cast<int>(lo - m.begin()), static_cast<int>(hi - lo)};
    }
    // |m0 - m| <= apm (zocrs/common.cu:14): every rounding step is monotone in m, so the
    // kept traces are contiguous in the sorted line
    auto inside = [&](float x) { return std::fabs(m0 - x) <= apm; };
    auto lo = std::partition_point(m.begin(), m.end(), [&](float x) { return x < m0 && !inside(x); });
    auto hi = std::partition_point(lo, m.end(), [&](float x) { return !(x > m0 && !inside(x)); });
    return {static_cast<int>(lo - m.begin()), static_cast<int>(hi - lo)};
}

// Index lists from device masks [n][ntr] (the OCT selections).
static std::vector<std::vector<int32_t>> masks_to_lists(const DeviceBuffer &mask, size_t n, size_t ntr,
                                                        hipStream_t stream) {
    std::vector<uint8_t> host(n * ntr);
    mask.download(host.data(), host.size(), stream);
    std::vector<std::vector<int32_t>> lists(n);
    for (size_t g = 0; g < n; g++) {
        for (size_t i = 0; i < ntr; i++) {
            if (host[g * ntr + i]) lists[g].push_back(static_cast<int32_t>(i));
        }
    }
    return lists;
}

// OCT selection (hss_select_traces) of the trace lists of m0s, in chunks of midpoints: at
// most 65535 per call (the selection kernel's grid limit) and about 256 MiB of per-trace
// mask each (the environment variable HSS_SELECT_MASK_BYTES sets another budget), so a
// whole line in one batch needs neither a mask of midpoints x traces bytes nor one call
// past the limit.  `chunk(first, oct)` adapts the candidate description to the chunk
// starting at midpoint `first` (per-sample parameters).
static size_t select_mask_bytes() {
    const char *v = std::getenv("HSS_SELECT_MASK_BYTES");
    if (v && *v) {
        const unsigned long long b = std::strtoull(v, nullptr, 10);
        if (b > 0) return static_cast<size_t>(b);
    }
    return size_t(256) << 20;
}

template <class Chunk>
static std::vector<std::vector<int32_t>> select_oct_lists(const HipLine &line, const std::vector<float> &m0s,
                                                          float apm, float h0, hss_oct_candidates oct,
                                                          DeviceBuffer &dMask, hipStream_t stream,
                                                          Chunk &&chunk) {
    const size_t ntr = std::max<size_t>(line.traces().ntr, 1);
    const size_t per = std::max<size_t>(1, std::min<size_t>(65535, select_mask_bytes() / ntr));
    std::vector<std::vector<int32_t>> lists;
    lists.reserve(m0s.size());
    DeviceBuffer dm0(std::max<size_t>(std::min(per, m0s.size()), 1) * 4);
    for (size_t first = 0; first < m0s.size(); first += per) {
        const size_t n = std::min(per, m0s.size() - first);
        dm0.upload(m0s.data() + first, n * 4, stream);
        dMask.resize(std::max<size_t>(n * ntr, 1));
        hss_oct_candidates c = oct;
        chunk(first, c);
        HSS_ASSERT(hss_select_traces(&line.traces(), OCT, static_cast<int32_t>(n), dm0.as<float>(), apm,
                                     h0, &c, dMask.as<uint8_t>(), stream));
        for (auto &l : masks_to_lists(dMask, n, line.traces().ntr, stream)) lists.push_back(std::move(l));
    }
    return lists;
}

// OCT trace lists for the non-exact variants are summed grouped by half offset: the
// kernels then keep the offset-dependent half of the OCT traveltime across each group
// (the order changes the sums' rounding only; the exact variant keeps the reference's
// order, oct/*.cu's trace loops over the selected (m, h) traces).
// The sort key is a total order on the offset's bit pattern (ascending for ordinary
// values), so equal patterns -- the kernels' test for one run -- sit together and a NaN
// offset cannot break the ordering.
static void group_by_halfoffset(std::vector<int32_t> &list, const HipLine &line) {
    if (BackendOptions::get().variant == HSS_EXACT) return;
    const std::vector<float> &h = line.halfoffsets();
    auto key = [&](int32_t i) {
        uint32_t u;
        std::memcpy(&u, &h[i], sizeof u);
        return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    };
    std::stable_sort(list.begin(), list.end(), [&](int32_t a, int32_t b) { return key(a) < key(b); });
}

// Launch order of a batch's midpoints.  The kernels hand (midpoint, tile) workgroups to
// the 8 XCDs in contiguous eighths of the launch (logical_block in the kernels), and the
// folds of a line differ; sorting the midpoints by work and dealing them to 8 bins in
// snake order (0..7, 7..0, ...) gives every XCD the same number of midpoints with about
// the same total work, each bin largest first.  Returns a permutation of [0, n).
std::vector<int32_t> xcd_balanced_order(const std::vector<int32_t> &work) {
    const size_t n = work.size(), bins = 8;
    std::vector<int32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return work[a] > work[b]; });
    std::vector<std::vector<int32_t>> bin(bins);
    const size_t full = n / bins;  // complete rounds; the last partial one fills bins 0.. so
                                   // that the larger bins come first, as the XCD split does
    for (size_t k = 0; k < n; k++) {
        const size_t round = k / bins, pos = k % bins;
        bin[(round & 1) && round < full ? bins - 1 - pos : pos].push_back(idx[k]);
    }
    std::vector<int32_t> order;
    order.reserve(n);
    for (auto &b : bin) order.insert(order.end(), b.begin(), b.end());
    return order;
}

// Statistics of one midpoint of a batch: its share of the batch's time, in proportion to
// its (selected traces x evaluations) work.
static std::array<float, kStatisticCount> batch_statistics(double work, double totalWork,
                                                           double notUsed, double kernelSec,
                                                           double selectionSec, unsigned selected,
                                                           unsigned window) {
    std::array<float, kStatisticCount> st{};
    const double share = totalWork > 0 ? work / totalWork : 0.0;
    const double k = kernelSec * share, sel = selectionSec * share;
    const double interp = static_cast<double>(window) * (work - notUsed);
    st[static_cast<int>(StatisticResult::INTR_PER_SEC)] = k > 0 ? static_cast<float>(interp / k) : 0.0f;
    st[static_cast<int>(StatisticResult::EFFICIENCY)] =
        static_cast<float>(work > 0 ? 1.0 - notUsed / work : 1.0);
    st[static_cast<int>(StatisticResult::SELECTED_TRACES)] = static_cast<float>(selected);
    st[static_cast<int>(StatisticResult::TOTAL_SELECTION_KERNEL_EXECUTION_TIME)] = static_cast<float>(sel);
    st[static_cast<int>(StatisticResult::TOTAL_KERNEL_EXECUTION_TIME)] = static_cast<float>(k);
    return st;
}

// Times work queued on `stream` with HIP events.
class StreamTimer {
  public:
    explicit StreamTimer(hipStream_t s) : stream(s) {
        HIP_ASSERT(hipEventCreate(&a));
        HIP_ASSERT(hipEventCreate(&b));
        HIP_ASSERT(hipEventRecord(a, stream));
    }
    ~StreamTimer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    double stopSeconds() {
        HIP_ASSERT(hipEventRecord(b, stream));
        HIP_ASSERT(hipEventSynchronize(b));
        float ms = 0;
        HIP_ASSERT(hipEventElapsedTime(&ms, a, b));
        return ms * 1e-3;
    }

  private:
    hipStream_t stream;
    hipEvent_t a{}, b{};
};

// ---------------------------------------------------------------------------------
hipStream_t HipLinearSearchAlgorithm::stream() const { return stream_of(deviceContext); }

void HipLinearSearchAlgorithm::copyGatherDataToDevice() {
    line.upload(*Gather::getInstance(), BackendOptions::get().storage, stream());
    dNotUsed.resize(sizeof(unsigned long long));
}

void HipLinearSearchAlgorithm::initializeParameters() {
    const unsigned np = traveltime->getNumberOfParameters();
    float lo[3] = {0, 0, 0}, st[3] = {0, 0, 0};
    int32_t cnt[3] = {1, 1, 1};
    for (unsigned p = 0; p < np; p++) {
        lo[p] = traveltime->getLowerBoundForParameter(p);
        st[p] = discretizationStep[p];
        cnt[p] = static_cast<int32_t>(discretizationGranularity[p]);
    }
    HSS_ASSERT(hss_build_parameter_array(traveltime->getModel(), lo, st, cnt,
                                         HIP_DEV_PTR(deviceParameterArray), stream()));
    if (traveltime->getModel() == OCT) {
        // the selection draws of CudaLinearSearchAlgorithm.cu:268-292: a default-constructed
        // engine, all velocities (as 4/v^2) first, then all slopes
        const unsigned samplePop = 1024;
        std::vector<float> pairs(samplePop * 2);
        std::default_random_engine generator;
        for (unsigned p = 0; p < 2; p++) {
            std::uniform_real_distribution<float> dist(traveltime->getLowerBoundForParameter(p),
                                                       traveltime->getUpperBoundForParameter(p));
            for (unsigned i = 0; i < samplePop; i++) {
                float r = dist(generator);
                if (p == OffsetContinuationTrajectory::VELOCITY) r = 4.0f / (r * r);
                pairs[i * 2 + p] = r;
            }
        }
        dOctPairs.resize(pairs.size() * sizeof(float));
        dOctPairs.upload(pairs.data(), pairs.size() * sizeof(float), stream());
    }
    HIP_ASSERT(hipStreamSynchronize(stream()));
}

void HipLinearSearchAlgorithm::selectBatch(const std::vector<float> &m0s) {
    const Gather *gather = Gather::getInstance();
    const traveltime_t model = traveltime->getModel();
    batch.clear();
    if (model != OCT) {
        for (float m0 : m0s) {
            const TraceRange r = selectRange(line.midpoints(), model, m0, gather->getApm());
            batch.addRange(m0, r.begin, r.count);
        }
    } else {
        hss_oct_candidates oct{};
        oct.mode = HSS_OCT_PAIRS;
        oct.d_pairs = dOctPairs.as<float>();
        oct.npairs = static_cast<int32_t>(dOctPairs.size() / (2 * sizeof(float)));
        auto lists = select_oct_lists(line, m0s, gather->getApm(), traveltime->getReferenceHalfoffset(), oct,
                                      dMask, stream(), [](size_t, hss_oct_candidates &) {});
        for (size_t g = 0; g < m0s.size(); g++) {
            group_by_halfoffset(lists[g], line);
            batch.addIndexList(m0s[g], lists[g]);
        }
    }
    batch.upload(stream());
}

void HipLinearSearchAlgorithm::selectTracesToBeUsedForMidpoint(float m0) {
    selectBatch({m0});
    filteredTracesCount = static_cast<unsigned>(batch.count(0));
    LOGH("Selected " << filteredTracesCount << " traces");
}

void HipLinearSearchAlgorithm::computeSemblanceAtGpuForMidpoint(float m0) {
    lastNotUsed = 0;
    if (!filteredTracesCount) {
        LOGI("midpoint " << m0 << ": no trace in the aperture, results stay zero");
        return;
    }
    HIP_ASSERT(hipMemsetAsync(dNotUsed.get(), 0, sizeof(unsigned long long), stream()));
    const hss_midpoints mids = batch.view(traveltime->getReferenceHalfoffset(), Gather::getInstance()->getApm());
    HSS_ASSERT(hss_grid_semblance(&line.traces(), &mids, traveltime->getModel(),
                                  HIP_DEV_PTR(deviceParameterArray), getTotalNumberOfParameters(),
                                  BackendOptions::get().variant, HIP_DEV_PTR(deviceResultArray),
                                  nullptr, nullptr, dNotUsed.as<unsigned long long>(), stream()));
    unsigned long long nu = 0;
    dNotUsed.download(&nu, sizeof(nu), stream());
    lastNotUsed = static_cast<float>(nu);
}

void HipLinearSearchAlgorithm::computeSemblanceAndParametersForMidpoints(
    const std::vector<float> &m0s, std::vector<MidpointOutput> &out) {
    const Gather *gather = Gather::getInstance();
    const unsigned ns = gather->getSamplesPerTrace();
    const unsigned nres = traveltime->getNumberOfResults();
    const unsigned P = getTotalNumberOfParameters();
    deviceContext->activate();
    auto t0 = std::chrono::steady_clock::now();
    selectBatch(m0s);
    const double selSec = Seconds(std::chrono::steady_clock::now() - t0).count();
    // midpoints without traces are skipped, their results stay zero (the reference returns
    // before launching, CudaLinearSearchAlgorithm.cu:37-40)
    std::vector<float> launchM0;
    std::vector<size_t> slot;
    for (size_t g = 0; g < m0s.size(); g++) {
        if (batch.count(g) > 0) {
            launchM0.push_back(m0s[g]);
            slot.push_back(g);
        }
    }
    std::vector<int> counts(m0s.size());
    for (size_t g = 0; g < m0s.size(); g++) counts[g] = batch.count(g);
    // one launch for all of them (the batch carries the XCD-balanced launch order)
    if (slot.size() != m0s.size()) {
        batch.retain(slot);
        batch.upload(stream());
    }
    out.assign(m0s.size(), MidpointOutput{});
    std::vector<unsigned long long> nu(launchM0.size(), 0);
    std::vector<float> res(static_cast<size_t>(launchM0.size()) * nres * ns, 0.0f);
    double kernelSec = 0;
    if (!launchM0.empty()) {
        DeviceBuffer dRes(res.size() * sizeof(float));
        DeviceBuffer dNu(nu.size() * sizeof(unsigned long long));
        HIP_ASSERT(hipMemsetAsync(dNu.get(), 0, dNu.size(), stream()));
        const hss_midpoints mids = batch.view(traveltime->getReferenceHalfoffset(), gather->getApm());
        StreamTimer timer(stream());
        HSS_ASSERT(hss_grid_semblance(&line.traces(), &mids, traveltime->getModel(),
                                      HIP_DEV_PTR(deviceParameterArray), P,
                                      BackendOptions::get().variant, dRes.as<float>(), nullptr,
                                      nullptr, dNu.as<unsigned long long>(), stream()));
        polish_results(line, mids, *traveltime, dRes, stream());
        kernelSec = timer.stopSeconds();
        dRes.download(res.data(), res.size() * sizeof(float), stream());
        dNu.download(nu.data(), nu.size() * sizeof(unsigned long long), stream());
    }
    double totalWork = 0;
    for (size_t g = 0; g < m0s.size(); g++) totalWork += static_cast<double>(counts[g]) * ns * P;
    std::vector<double> notUsed(m0s.size(), 0.0);
    for (size_t j = 0; j < slot.size(); j++) notUsed[slot[j]] = static_cast<double>(nu[j]);
    for (size_t g = 0; g < m0s.size(); g++) {
        out[g].m0 = m0s[g];
        out[g].results.assign(static_cast<size_t>(nres) * ns, 0.0f);
    }
    for (size_t j = 0; j < slot.size(); j++) {
        std::copy(res.begin() + static_cast<std::ptrdiff_t>(j * nres * ns),
                  res.begin() + static_cast<std::ptrdiff_t>((j + 1) * nres * ns),
                  out[slot[j]].results.begin());
    }
    for (size_t g = 0; g < m0s.size(); g++) {
        out[g].statistics = batch_statistics(static_cast<double>(counts[g]) * ns * P, totalWork,
                                             notUsed[g], kernelSec, selSec,
                                             static_cast<unsigned>(counts[g]), gather->getWindowSize());
    }
    LOGI("Computed " << m0s.size() << " midpoints in one launch: " << kernelSec << " s");
}

// ---------------------------------------------------------------------------------
hipStream_t HipDifferentialEvolutionAlgorithm::stream() const { return stream_of(deviceContext); }

void HipDifferentialEvolutionAlgorithm::copyGatherDataToDevice() {
    const Gather *gather = Gather::getInstance();
    line.upload(*gather, BackendOptions::get().storage, stream());
    dNotUsed.resize(sizeof(unsigned long long));
    uint32_t k = 0;
    cdpOrdinal.clear();
    for (const auto &kv : gather->getCdps()) cdpOrdinal[kv.first] = k++;
}

void HipDifferentialEvolutionAlgorithm::setupRandomSeedArray() {
    BackendOptions &o = BackendOptions::get();
    // the reference seeds from the clock (srand(time(NULL)), CudaDifferentialEvolutionAlgorithm.cu:197-199)
    seed = o.seedGiven ? o.seed : static_cast<uint64_t>(std::time(nullptr));
    LOGI("Differential evolution seed = " << seed);
}

uint32_t HipDifferentialEvolutionAlgorithm::streamOf(float m0) const {
    auto it = cdpOrdinal.find(m0);
    return it == cdpOrdinal.end() ? 0u : it->second;
}

void polish_results(const HipLine &line, const hss_midpoints &mids, const Traveltime &tt,
                    DeviceBuffer &dRes, hipStream_t stream) {
    const unsigned steps = BackendOptions::get().polishSteps;
    if (steps == 0 || mids.n == 0) return;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (unsigned p = 0; p < tt.getNumberOfParameters() && p < 3; p++) {
        lo[p] = tt.getLowerBoundForParameter(p);
        hi[p] = tt.getUpperBoundForParameter(p);
    }
    DeviceBuffer dPol(dRes.size());
    HSS_ASSERT(hss_polish(&line.traces(), &mids, tt.getModel(), lo, hi, static_cast<int32_t>(steps),
                          dRes.as<float>(), dPol.as<float>(), stream));
    HIP_ASSERT(hipMemcpyAsync(dRes.get(), dPol.get(), dRes.size(), hipMemcpyDeviceToDevice, stream));
    HIP_ASSERT(hipStreamSynchronize(stream));
}

void HipDifferentialEvolutionAlgorithm::bounds(float lo[3], float hi[3]) const {
    for (unsigned p = 0; p < 3; p++) {
        const bool has = p < traveltime->getNumberOfParameters();
        lo[p] = has ? traveltime->getLowerBoundForParameter(p) : 0.0f;
        hi[p] = has ? traveltime->getUpperBoundForParameter(p) : 0.0f;
    }
}

void HipDifferentialEvolutionAlgorithm::selectBatch(const std::vector<float> &m0s,
                                                    const std::vector<uint32_t> &streams) {
    const Gather *gather = Gather::getInstance();
    const traveltime_t model = traveltime->getModel();
    batch.clear();
    if (model != OCT) {
        for (float m0 : m0s) {
            const TraceRange r = selectRange(line.midpoints(), model, m0, gather->getApm());
            batch.addRange(m0, r.begin, r.count);
        }
    } else {
        // the start population of each midpoint (oct/differential_evolution.cu:130-168)
        const size_t ntr = line.traces().ntr;
        dMask.resize(std::max<size_t>(ntr, 1));
        DeviceBuffer dm0(4);
        float lo[3], hi[3];
        bounds(lo, hi);
        for (size_t g = 0; g < m0s.size(); g++) {
            dm0.upload(&m0s[g], 4, stream());
            hss_oct_candidates oct{};
            oct.mode = HSS_OCT_DE_START;
            oct.NP = static_cast<int32_t>(individualsPerPopulation);
            oct.lower[0] = lo[0];
            oct.lower[1] = lo[1];
            oct.upper[0] = hi[0];
            oct.upper[1] = hi[1];
            oct.seed = seed;
            oct.stream_base = streams[g];
            HSS_ASSERT(hss_select_traces(&line.traces(), model, 1, dm0.as<float>(), gather->getApm(),
                                         traveltime->getReferenceHalfoffset(), &oct,
                                         dMask.as<uint8_t>(), stream()));
            std::vector<int32_t> list = masks_to_lists(dMask, 1, ntr, stream())[0];
            group_by_halfoffset(list, line);
            batch.addIndexList(m0s[g], list);
        }
    }
    batch.upload(stream());
}

void HipDifferentialEvolutionAlgorithm::selectTracesToBeUsedForMidpoint(float m0) {
    selectBatch({m0}, {streamOf(m0)});
    filteredTracesCount = static_cast<unsigned>(batch.count(0));
    LOGH("Selected " << filteredTracesCount << " traces");
}

void HipDifferentialEvolutionAlgorithm::runGenerations(const hss_midpoints &mids, uint32_t streamBase,
                                                       int g0, int g1, float *dResult) {
    uint64_t need = 0;
    HSS_ASSERT(hss_de_workspace_size(mids.n, line.traces().ns, traveltime->getModel(),
                                     static_cast<int32_t>(individualsPerPopulation), &need));
    if (dWorkspace.size() < need) dWorkspace.resize(need);
    float lo[3], hi[3];
    bounds(lo, hi);
    HSS_ASSERT(hss_de_semblance(&line.traces(), &mids, traveltime->getModel(), lo, hi,
                                static_cast<int32_t>(individualsPerPopulation),
                                static_cast<int32_t>(generations), g0, g1, seed, streamBase,
                                BackendOptions::get().variant, dWorkspace.get(), dWorkspace.size(),
                                dResult, nullptr, nullptr, dNotUsed.as<unsigned long long>(),
                                stream()));
}

void HipDifferentialEvolutionAlgorithm::computeSemblanceAtGpuForMidpoint(float m0) {
    // one generation of the stepwise interface
    if (nextGeneration == 0) HIP_ASSERT(hipMemsetAsync(dNotUsed.get(), 0, sizeof(unsigned long long), stream()));
    const hss_midpoints mids = batch.view(traveltime->getReferenceHalfoffset(), Gather::getInstance()->getApm());
    runGenerations(mids, streamOf(m0), nextGeneration, nextGeneration + 1, HIP_DEV_PTR(deviceResultArray));
    nextGeneration++;
    unsigned long long nu = 0;
    dNotUsed.download(&nu, sizeof(nu), stream());
    lastNotUsed = static_cast<float>(nu);
}

void HipDifferentialEvolutionAlgorithm::selectBestIndividuals(std::vector<float> &resultArrays) {
    deviceResultArray->pasteTo(resultArrays);
}

void HipDifferentialEvolutionAlgorithm::computeSemblanceAndParametersForMidpoint(float m0) {
    LOGI("Computing semblance for m0 = " << m0);
    const unsigned ns = Gather::getInstance()->getSamplesPerTrace();
    deviceContext->activate();
    Seconds selection = Seconds::zero(), total = Seconds::zero();
    MEASURE_EXEC_TIME(selection, selectTracesToBeUsedForMidpoint(m0));
    HIP_ASSERT(hipMemsetAsync(dNotUsed.get(), 0, sizeof(unsigned long long), stream()));
    const hss_midpoints mids = batch.view(traveltime->getReferenceHalfoffset(), Gather::getInstance()->getApm());
    auto t0 = std::chrono::steady_clock::now();
    runGenerations(mids, streamOf(m0), 0, static_cast<int>(generations), HIP_DEV_PTR(deviceResultArray));
    HIP_ASSERT(hipStreamSynchronize(stream()));
    total += std::chrono::steady_clock::now() - t0;
    deviceResultArray->pasteTo(computedResults);
    unsigned long long nu = 0;
    dNotUsed.download(&nu, sizeof(nu), stream());
    lastNotUsed = static_cast<float>(nu);
    const float used = static_cast<float>(filteredTracesCount) * static_cast<float>(ns) *
                       static_cast<float>(generations) * static_cast<float>(individualsPerPopulation);
    saveStatisticalResults(used, total, selection);
}

void HipDifferentialEvolutionAlgorithm::computeSemblanceAndParametersForMidpoints(
    const std::vector<float> &m0s, std::vector<MidpointOutput> &out) {
    // one launch needs consecutive stream ids: split the batch into runs of consecutive CDPs
    out.clear();
    size_t start = 0;
    const Gather *gather = Gather::getInstance();
    const unsigned ns = gather->getSamplesPerTrace();
    const unsigned nres = traveltime->getNumberOfResults();
    deviceContext->activate();
    while (start < m0s.size()) {
        size_t end = start + 1;
        while (end < m0s.size() && streamOf(m0s[end]) == streamOf(m0s[start]) + (end - start)) end++;
        const std::vector<float> run(m0s.begin() + static_cast<std::ptrdiff_t>(start),
                                     m0s.begin() + static_cast<std::ptrdiff_t>(end));
        std::vector<uint32_t> streams(run.size());
        for (size_t g = 0; g < run.size(); g++) streams[g] = streamOf(run[g]);
        auto t0 = std::chrono::steady_clock::now();
        selectBatch(run, streams);
        const double selSec = Seconds(std::chrono::steady_clock::now() - t0).count();
        std::vector<float> res(run.size() * nres * ns);
        DeviceBuffer dRes(res.size() * sizeof(float));
        DeviceBuffer dNu(run.size() * sizeof(unsigned long long));
        HIP_ASSERT(hipMemsetAsync(dNu.get(), 0, dNu.size(), stream()));
        const hss_midpoints mids = batch.view(traveltime->getReferenceHalfoffset(), gather->getApm());
        uint64_t need = 0;
        HSS_ASSERT(hss_de_workspace_size(mids.n, ns, traveltime->getModel(),
                                         static_cast<int32_t>(individualsPerPopulation), &need));
        if (dWorkspace.size() < need) dWorkspace.resize(need);
        float lo[3], hi[3];
        bounds(lo, hi);
        StreamTimer timer(stream());
        HSS_ASSERT(hss_de_semblance(&line.traces(), &mids, traveltime->getModel(), lo, hi,
                                    static_cast<int32_t>(individualsPerPopulation),
                                    static_cast<int32_t>(generations), 0,
                                    static_cast<int32_t>(generations), seed, streams[0],
                                    BackendOptions::get().variant, dWorkspace.get(),
                                    dWorkspace.size(), dRes.as<float>(), nullptr, nullptr,
                                    dNu.as<unsigned long long>(), stream()));
        polish_results(line, mids, *traveltime, dRes, stream());
        const double kernelSec = timer.stopSeconds();
        std::vector<unsigned long long> nu(run.size());
        dRes.download(res.data(), res.size() * sizeof(float), stream());
        dNu.download(nu.data(), nu.size() * sizeof(unsigned long long), stream());
        double totalWork = 0;
        for (size_t g = 0; g < run.size(); g++)
            totalWork += static_cast<double>(batch.count(g)) * ns * generations * individualsPerPopulation;
        for (size_t g = 0; g < run.size(); g++) {
            MidpointOutput o;
            o.m0 = run[g];
            o.results.assign(res.begin() + static_cast<std::ptrdiff_t>(g * nres * ns),
                             res.begin() + static_cast<std::ptrdiff_t>((g + 1) * nres * ns));
            const double work = static_cast<double>(batch.count(g)) * ns * generations * individualsPerPopulation;
            o.statistics = batch_statistics(work, totalWork, static_cast<double>(nu[g]), kernelSec, selSec,
                                            static_cast<unsigned>(batch.count(g)), gather->getWindowSize());
            out.push_back(std::move(o));
        }
        LOGI("Evolved " << run.size() << " midpoints in one launch: " << kernelSec << " s");
        start = end;
    }
}

// ---------------------------------------------------------------------------------
hipStream_t HipStretchFreeAlgorithm::stream() const { return stream_of(deviceContext); }

void HipStretchFreeAlgorithm::copyGatherDataToDevice() {
    line.upload(*Gather::getInstance(), BackendOptions::get().storage, stream());
    dNotUsed.resize(sizeof(unsigned long long));
}

void HipStretchFreeAlgorithm::uploadParameters(const std::vector<float> &m0s) {
    const unsigned ns = Gather::getInstance()->getSamplesPerTrace();
    const unsigned np = traveltime->getNumberOfParameters();
    std::vector<float> v(m0s.size() * np * ns);
    for (size_t g = 0; g < m0s.size(); g++) {
        const auto &p = nonStretchFreeParameters.at(m0s[g]);
        std::copy(p.begin(), p.end(), v.begin() + static_cast<std::ptrdiff_t>(g * np * ns));
    }
    dVparams.resize(std::max<size_t>(v.size(), 1) * sizeof(float));
    dVparams.upload(v.data(), v.size() * sizeof(float), stream());
}

void HipStretchFreeAlgorithm::selectBatch(const std::vector<float> &m0s) {
    const Gather *gather = Gather::getInstance();
    const traveltime_t model = traveltime->getModel();
    uploadParameters(m0s);
    batch.clear();
    if (model != OCT) {
        for (float m0 : m0s) {
            const TraceRange r = selectRange(line.midpoints(), model, m0, gather->getApm());
            batch.addRange(m0, r.begin, r.count);
        }
    } else {
        // the per-sample (v, slope) of each midpoint (oct/stretch_free.cu:99-138)
        hss_oct_candidates oct{};
        oct.mode = HSS_OCT_PER_SAMPLE;
        const float *vparams = dVparams.as<float>();
        const size_t perMidpoint = size_t(2) * line.traces().ns;  // [n][2][ns]
        oct.d_vparams = vparams;
        auto lists = select_oct_lists(line, m0s, gather->getApm(), traveltime->getReferenceHalfoffset(), oct,
                                      dMask, stream(), [&](size_t first, hss_oct_candidates &c) {
                                          c.d_vparams = vparams + first * perMidpoint;
                                      });
        for (size_t g = 0; g < m0s.size(); g++) {
            group_by_halfoffset(lists[g], line);
            batch.addIndexList(m0s[g], lists[g]);
        }
    }
    batch.upload(stream());
}

void HipStretchFreeAlgorithm::selectTracesToBeUsedForMidpoint(float m0) {
    selectBatch({m0});
    filteredTracesCount = static_cast<unsigned>(batch.count(0));
    LOGH("Selected " << filteredTracesCount << " traces");
}

void HipStretchFreeAlgorithm::computeSemblanceAtGpuForMidpoint(float m0) {
    lastNotUsed = 0;
    if (!filteredTracesCount) {
        LOGI("midpoint " << m0 << ": no trace in the aperture, results stay zero");
        return;
    }
    HIP_ASSERT(hipMemsetAsync(dNotUsed.get(), 0, sizeof(unsigned long long), stream()));
    const hss_midpoints mids = batch.view(traveltime->getReferenceHalfoffset(), Gather::getInstance()->getApm());
    HSS_ASSERT(hss_stretch_free(&line.traces(), &mids, traveltime->getModel(), dVparams.as<float>(),
                                nArray.data(), static_cast<int32_t>(nArray.size()),
                                BackendOptions::get().variant, HIP_DEV_PTR(deviceResultArray), nullptr,
                                nullptr, dNotUsed.as<unsigned long long>(), stream()));
    unsigned long long nu = 0;
    dNotUsed.download(&nu, sizeof(nu), stream());
    lastNotUsed = static_cast<float>(nu);
}

void HipStretchFreeAlgorithm::computeSemblanceAndParametersForMidpoints(
    const std::vector<float> &m0s, std::vector<MidpointOutput> &out) {
    const Gather *gather = Gather::getInstance();
    const unsigned ns = gather->getSamplesPerTrace();
    const unsigned nres = traveltime->getNumberOfResults();
    const unsigned nN = static_cast<unsigned>(nArray.size());
    deviceContext->activate();
    auto t0 = std::chrono::steady_clock::now();
    selectBatch(m0s);
    std::vector<float> launchM0;
    std::vector<size_t> slot;
    std::vector<int> counts(m0s.size());
    for (size_t g = 0; g < m0s.size(); g++) {
        counts[g] = batch.count(g);
        if (counts[g] > 0) {
            launchM0.push_back(m0s[g]);
            slot.push_back(g);
        }
    }
    if (launchM0.size() != m0s.size()) selectBatch(launchM0);
    const double selSec = Seconds(std::chrono::steady_clock::now() - t0).count();
    out.assign(m0s.size(), MidpointOutput{});
    std::vector<float> res(launchM0.size() * nres * ns, 0.0f);
    std::vector<unsigned long long> nu(launchM0.size(), 0);
    double kernelSec = 0;
    if (!launchM0.empty()) {
        DeviceBuffer dRes(res.size() * sizeof(float));
        DeviceBuffer dNu(nu.size() * sizeof(unsigned long long));
        HIP_ASSERT(hipMemsetAsync(dNu.get(), 0, dNu.size(), stream()));
        const hss_midpoints mids = batch.view(traveltime->getReferenceHalfoffset(), gather->getApm());
        StreamTimer timer(stream());
        HSS_ASSERT(hss_stretch_free(&line.traces(), &mids, traveltime->getModel(), dVparams.as<float>(),
                                    nArray.data(), static_cast<int32_t>(nN),
                                    BackendOptions::get().variant, dRes.as<float>(), nullptr, nullptr,
                                    dNu.as<unsigned long long>(), stream()));
        kernelSec = timer.stopSeconds();
        dRes.download(res.data(), res.size() * sizeof(float), stream());
        dNu.download(nu.data(), nu.size() * sizeof(unsigned long long), stream());
    }
    double totalWork = 0;
    for (size_t g = 0; g < m0s.size(); g++) totalWork += static_cast<double>(counts[g]) * ns * nN;
    std::vector<double> notUsed(m0s.size(), 0.0);
    for (size_t j = 0; j < slot.size(); j++) notUsed[slot[j]] = static_cast<double>(nu[j]);
    for (size_t g = 0; g < m0s.size(); g++) {
        out[g].m0 = m0s[g];
        out[g].results.assign(static_cast<size_t>(nres) * ns, 0.0f);
        out[g].statistics = batch_statistics(static_cast<double>(counts[g]) * ns * nN, totalWork,
                                             notUsed[g], kernelSec, selSec,
                                             static_cast<unsigned>(counts[g]), gather->getWindowSize());
    }
    for (size_t j = 0; j < slot.size(); j++) {
        std::copy(res.begin() + static_cast<std::ptrdiff_t>(j * nres * ns),
                  res.begin() + static_cast<std::ptrdiff_t>((j + 1) * nres * ns),
                  out[slot[j]].results.begin());
    }
}

// ---------------------------------------------------------------------------------
ComputeAlgorithmBuilder *HipComputeAlgorithmBuilder::getInstance() {
    static HipComputeAlgorithmBuilder builder;
    return &builder;
}

LinearSearchAlgorithm *HipComputeAlgorithmBuilder::buildLinearSearchAlgorithm(
    std::shared_ptr<Traveltime> traveltime, std::shared_ptr<DeviceContext> context,
    unsigned int threadCount, const std::vector<int> &discretization) {
    if (discretization.size() != traveltime->getNumberOfParameters()) {
        std::ostringstream e;
        e << traveltime->getTraveltimeWord() << " requires exact "
          << traveltime->getNumberOfParameters() << " discretization granularities.";
        throw std::logic_error(e.str());
    }
    auto *a = new HipLinearSearchAlgorithm(traveltime, context, HipDataContainerBuilder::getInstance(),
                                           threadCount);
    for (unsigned i = 0; i < traveltime->getNumberOfParameters(); i++) {
        a->setDiscretizationGranularityForParameter(i, static_cast<unsigned>(discretization[i]));
    }
    return a;
}

DifferentialEvolutionAlgorithm *HipComputeAlgorithmBuilder::buildDifferentialEvolutionAlgorithm(
    std::shared_ptr<Traveltime> traveltime, std::shared_ptr<DeviceContext> context,
    unsigned int threadCount, unsigned int generation, unsigned int individuals) {
    return new HipDifferentialEvolutionAlgorithm(traveltime, context, HipDataContainerBuilder::getInstance(),
                                                 threadCount, generation, individuals);
}

StretchFreeAlgorithm *HipComputeAlgorithmBuilder::buildStretchFreeAlgorithm(
    std::shared_ptr<Traveltime> traveltime, std::shared_ptr<DeviceContext> context,
    unsigned int threadCount, const std::vector<std::string> &files) {
    if (files.size() != traveltime->getNumberOfParameters()) {
        std::ostringstream e;
        e << traveltime->getTraveltimeWord() << " requires exact "
          << traveltime->getNumberOfParameters() << " parameter files.";
        throw std::logic_error(e.str());
    }
    return new HipStretchFreeAlgorithm(traveltime, context, HipDataContainerBuilder::getInstance(),
                                       threadCount, files);
}
