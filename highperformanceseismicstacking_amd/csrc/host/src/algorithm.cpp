// Common per-midpoint drivers (see algorithm.hpp).
#include "algorithm.hpp"

#include <fstream>
#include <numeric>
#include <sstream>
#include <stdexcept>

#include "logger.hpp"

ComputeAlgorithm::ComputeAlgorithm(const std::string &name, std::shared_ptr<Traveltime> tt,
                                   std::shared_ptr<DeviceContext> context,
                                   DataContainerBuilder *dataBuilder, unsigned int threads)
    : dataFactory(dataBuilder), algorithmName(name), deviceContext(std::move(context)),
      traveltime(std::move(tt)), threadCount(threads) {}

float ComputeAlgorithm::getStatisticalResult(StatisticResult r) const {
    auto it = computedStatisticalResults.find(static_cast<int>(r));
    if (it == computedStatisticalResults.end()) {
        throw std::invalid_argument("no value recorded for this statistic");
    }
    return it->second;
}

void ComputeAlgorithm::saveStatisticalResults(float totalUsedTracesCount, Seconds totalTime,
                                              Seconds selectionTime) {
    const float notUsed = getNotUsedCount();
    const float interpolations =
        static_cast<float>(Gather::getInstance()->getWindowSize()) * (totalUsedTracesCount - notUsed);
    auto &st = computedStatisticalResults;
    st[static_cast<int>(StatisticResult::INTR_PER_SEC)] =
        interpolations / static_cast<float>(totalTime.count());
    float efficiency = 1.0f;
    if (totalUsedTracesCount > 0) efficiency -= notUsed / totalUsedTracesCount;
    st[static_cast<int>(StatisticResult::EFFICIENCY)] = efficiency;
    st[static_cast<int>(StatisticResult::SELECTED_TRACES)] = static_cast<float>(filteredTracesCount);
    st[static_cast<int>(StatisticResult::TOTAL_SELECTION_KERNEL_EXECUTION_TIME)] =
        static_cast<float>(selectionTime.count());
    st[static_cast<int>(StatisticResult::TOTAL_KERNEL_EXECUTION_TIME)] =
        static_cast<float>(totalTime.count());
    LOGD("trace selection: " << selectionTime.count() << " s in all");
    LOGI("semblance kernels: " << totalTime.count() << " s in all");
}

void ComputeAlgorithm::compileKernels(const std::string &path) {
    LOGD("Kernel path " << path << " ignored: HIP kernels are compiled ahead of time");
}

void ComputeAlgorithm::computeSemblanceAndParametersForMidpoints(const std::vector<float> &m0s,
                                                                 std::vector<MidpointOutput> &out) {
    out.clear();
    out.reserve(m0s.size());
    for (float m0 : m0s) {
        computeSemblanceAndParametersForMidpoint(m0);
        MidpointOutput o;
        o.m0 = m0;
        o.results = computedResults;
        for (unsigned s = 0; s < kStatisticCount; s++) {
            o.statistics[s] = getStatisticalResult(static_cast<StatisticResult>(s));
        }
        out.push_back(std::move(o));
    }
}

// ---------------------------------------------------------------------------------
LinearSearchAlgorithm::LinearSearchAlgorithm(std::shared_ptr<Traveltime> tt,
                                             std::shared_ptr<DeviceContext> context,
                                             DataContainerBuilder *dataBuilder, unsigned int threads)
    : ComputeAlgorithm("linear-search", tt, context, dataBuilder, threads),
      discretizationGranularity(tt->getNumberOfParameters()),
      discretizationDivisor(tt->getNumberOfParameters()),
      discretizationStep(tt->getNumberOfParameters()) {}

unsigned int LinearSearchAlgorithm::getTotalNumberOfParameters() const {
    return std::accumulate(discretizationGranularity.begin(), discretizationGranularity.end(), 1u,
                           std::multiplies<unsigned int>());
}

float LinearSearchAlgorithm::getParameterValueAt(unsigned int iteration, unsigned int p) const {
    const unsigned int step = (iteration / discretizationDivisor[p]) % discretizationGranularity[p];
    return traveltime->getLowerBoundForParameter(p) + static_cast<float>(step) * discretizationStep[p];
}

void LinearSearchAlgorithm::setDiscretizationDivisorForParameter(unsigned int p, unsigned int d) {
    if (p >= traveltime->getNumberOfParameters()) throw std::invalid_argument("parameter index beyond the model's parameter count");
    discretizationDivisor[p] = d;
}

void LinearSearchAlgorithm::setDiscretizationGranularityForParameter(unsigned int p, unsigned int g) {
    if (p >= traveltime->getNumberOfParameters()) throw std::invalid_argument("parameter index beyond the model's parameter count");
    discretizationGranularity[p] = g;
}

void LinearSearchAlgorithm::setupArrays() {
    const Gather *gather = Gather::getInstance();
    const unsigned int np = traveltime->getNumberOfParameters();
    const unsigned int nres = traveltime->getNumberOfResults();
    const unsigned int ns = gather->getSamplesPerTrace();
    deviceParameterArray.reset(dataFactory->build(np * getTotalNumberOfParameters(), deviceContext));
    deviceResultArray.reset(dataFactory->build(nres * ns, deviceContext));
    computedResults.assign(static_cast<size_t>(nres) * ns, 0.0f);
}

void LinearSearchAlgorithm::setupDiscretizationSteps() {
    unsigned int divisor = 1;
    for (unsigned int i = 0; i < traveltime->getNumberOfParameters(); i++) {
        const float lo = traveltime->getLowerBoundForParameter(i);
        const float hi = traveltime->getUpperBoundForParameter(i);
        discretizationDivisor[i] = divisor;
        divisor *= discretizationGranularity[i];
        discretizationStep[i] = (hi - lo) / static_cast<float>(discretizationGranularity[i]);
    }
}

void LinearSearchAlgorithm::setUp() {
    Seconds init = Seconds::zero();
    deviceContext->activate();
    setupArrays();
    copyGatherDataToDevice();
    setupDiscretizationSteps();
    MEASURE_EXEC_TIME(init, initializeParameters());
    LOGD("parameter grid built in " << init.count() << " s");
    isSet = true;
}

void LinearSearchAlgorithm::computeSemblanceAndParametersForMidpoint(float m0) {
    LOGI("midpoint " << m0 << ": evaluating");
    const unsigned int ns = Gather::getInstance()->getSamplesPerTrace();
    deviceContext->activate();
    deviceResultArray->reset();
    Seconds selection = Seconds::zero(), total = Seconds::zero();
    MEASURE_EXEC_TIME(selection, selectTracesToBeUsedForMidpoint(m0));
    MEASURE_EXEC_TIME(total, computeSemblanceAtGpuForMidpoint(m0));
    deviceResultArray->pasteTo(computedResults);
    const float used = static_cast<float>(filteredTracesCount) * static_cast<float>(ns) *
                       static_cast<float>(getTotalNumberOfParameters());
    saveStatisticalResults(used, total, selection);
}

const std::string LinearSearchAlgorithm::toString() const {
    std::ostringstream s;
    s << "Total number of parameters = " << getTotalNumberOfParameters() << std::endl;
    for (unsigned int i = 0; i < traveltime->getNumberOfParameters(); i++) {
        s << "# of discrete values for " << traveltime->getDescriptionForParameter(i) << " = "
          << discretizationGranularity[i] << std::endl;
    }
    s << "Thread count = " << threadCount << std::endl;
    return s.str();
}

// ---------------------------------------------------------------------------------
DifferentialEvolutionAlgorithm::DifferentialEvolutionAlgorithm(
    std::shared_ptr<Traveltime> tt, std::shared_ptr<DeviceContext> context,
    DataContainerBuilder *dataBuilder, unsigned int threads, unsigned int gen, unsigned int ind)
    : ComputeAlgorithm("de", tt, context, dataBuilder, threads), generations(gen),
      individualsPerPopulation(ind) {}

void DifferentialEvolutionAlgorithm::setUp() {
    const unsigned int ns = Gather::getInstance()->getSamplesPerTrace();
    const unsigned int nres = traveltime->getNumberOfResults();
    deviceContext->activate();
    copyGatherDataToDevice();
    deviceResultArray.reset(dataFactory->build(nres * ns, deviceContext));
    computedResults.assign(static_cast<size_t>(nres) * ns, 0.0f);
    setupRandomSeedArray();
    isSet = true;
}

void DifferentialEvolutionAlgorithm::computeSemblanceAndParametersForMidpoint(float m0) {
    LOGI("midpoint " << m0 << ": evaluating");
    const unsigned int ns = Gather::getInstance()->getSamplesPerTrace();
    Seconds selection = Seconds::zero(), total = Seconds::zero();
    deviceContext->activate();
    deviceResultArray->reset();
    startAllPopulations();
    // the selection may depend on the start population (OCT)
    MEASURE_EXEC_TIME(selection, selectTracesToBeUsedForMidpoint(m0));
    MEASURE_EXEC_TIME(total, computeSemblanceAtGpuForMidpoint(m0));
    for (unsigned int gen = 1; gen < generations; gen++) {
        mutateAllPopulations();
        crossoverPopulationIndividuals();
        MEASURE_EXEC_TIME(total, computeSemblanceAtGpuForMidpoint(m0));
        advanceGeneration();
    }
    selectBestIndividuals(computedResults);
    const float used = static_cast<float>(filteredTracesCount) * static_cast<float>(ns) *
                       static_cast<float>(generations) * static_cast<float>(individualsPerPopulation);
    saveStatisticalResults(used, total, selection);
}

const std::string DifferentialEvolutionAlgorithm::toString() const {
    std::ostringstream s;
    s << "Total number of generations = " << generations << std::endl;
    s << "Total number of ind. per population = " << individualsPerPopulation << std::endl;
    return s.str();
}

// ---------------------------------------------------------------------------------
StretchFreeAlgorithm::StretchFreeAlgorithm(std::shared_ptr<Traveltime> tt,
                                           std::shared_ptr<DeviceContext> context,
                                           DataContainerBuilder *dataBuilder, unsigned int threads,
                                           const std::vector<std::string> &files)
    : ComputeAlgorithm("strecth-free", tt, context, dataBuilder, threads), parameterFileArray(files) {}

void StretchFreeAlgorithm::setUp() {
    const unsigned int ns = Gather::getInstance()->getSamplesPerTrace();
    const unsigned int nres = traveltime->getNumberOfResults();
    deviceContext->activate();
    copyGatherDataToDevice();
    deviceResultArray.reset(dataFactory->build(nres * ns, deviceContext));
    computedResults.assign(static_cast<size_t>(nres) * ns, 0.0f);
    readNonStretchedFreeParameterFromFile();
    nArray.clear();
    for (int n = -NMAX; n <= NMAX; n++) nArray.push_back(static_cast<float>(n));
    deviceParameterArray.reset(dataFactory->build(static_cast<unsigned int>(nArray.size()), deviceContext));
    deviceParameterArray->copyFrom(nArray);
    isSet = true;
}

void StretchFreeAlgorithm::readNonStretchedFreeParameterFromFile() {
    const Gather *gather = Gather::getInstance();
    const unsigned int np = traveltime->getNumberOfParameters();
    const unsigned int ns = gather->getSamplesPerTrace();
    if (parameterFileArray.size() < np) {
        throw std::logic_error("one parameter file per traveltime parameter is required");
    }
    std::vector<std::ifstream> files(np);
    for (unsigned int i = 0; i < np; i++) {
        files[i].open(parameterFileArray[i], std::ios::binary);
        if (!files[i].is_open()) throw std::invalid_argument("cannot open parameter file " + parameterFileArray[i]);
    }
    for (const auto &kv : gather->getCdps()) {
        std::vector<float> v(static_cast<size_t>(np) * ns, 0.0f);
        for (unsigned int i = 0; i < np; i++) {
            Trace t;
            if (!t.read(files[i], gather->getAzimuthInRad(), trace_info_t{}) || t.getSampleCount() != ns) {
                throw std::runtime_error(parameterFileArray[i] + ": no trace of " + std::to_string(ns) +
                                         " samples for the CDP at m = " + std::to_string(kv.first));
            }
            std::copy(t.getSamples().begin(), t.getSamples().end(), v.begin() + static_cast<std::ptrdiff_t>(i) * ns);
        }
        nonStretchFreeParameters[kv.first] = std::move(v);
    }
}

void StretchFreeAlgorithm::computeSemblanceAndParametersForMidpoint(float m0) {
    LOGI("midpoint " << m0 << ": evaluating");
    const unsigned int ns = Gather::getInstance()->getSamplesPerTrace();
    deviceContext->activate();
    deviceResultArray->reset();
    Seconds selection = Seconds::zero(), total = Seconds::zero();
    MEASURE_EXEC_TIME(selection, selectTracesToBeUsedForMidpoint(m0));
    MEASURE_EXEC_TIME(total, computeSemblanceAtGpuForMidpoint(m0));
    deviceResultArray->pasteTo(computedResults);
    const float used = static_cast<float>(filteredTracesCount) * static_cast<float>(ns) *
                       static_cast<float>(getTotalNumberOfParameters());
    saveStatisticalResults(used, total, selection);
}

const std::string StretchFreeAlgorithm::toString() const {
    std::ostringstream s;
    s << "Total number of parameters = " << getTotalNumberOfParameters() << std::endl;
    s << "N in [ " << -NMAX << "," << NMAX << "]" << std::endl;
    s << "Parameter files: " << std::endl;
    for (const auto &f : parameterFileArray) s << f << std::endl;
    return s.str();
}
