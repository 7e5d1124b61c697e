#include "hip/include/semblance/algorithm/HipDifferentialEvolutionAlgorithm.hpp"
#include "hip/include/execution/HipUtils.hpp"

#include <cstdlib>
#include <ctime>

HipDifferentialEvolutionAlgorithm::HipDifferentialEvolutionAlgorithm(
    shared_ptr<Traveltime> traveltime,
    shared_ptr<DeviceContext> context,
    DataContainerBuilder* dataBuilder,
    unsigned int threadCount,
    unsigned int generation,
    unsigned int individualsPerPopulation
) : DifferentialEvolutionAlgorithm(traveltime, context, dataBuilder, threadCount, generation,
                                   individualsPerPopulation) {
}

HipDifferentialEvolutionAlgorithm::~HipDifferentialEvolutionAlgorithm() {
    if (workspace) (void)hipFree(workspace);
    if (result) (void)hipFree(result);
}

// The reference seeds curand with time(NULL) (CudaDifferentialEvolutionAlgorithm.cu:
// 197-199); here the Philox key is HSS_DE_SEED when set, else the clock.  Midpoint k of
// this algorithm uses Philox stream k.  setUp() calls this last, once the gather is read.
void HipDifferentialEvolutionAlgorithm::setupRandomSeedArray() {
    const char* s = std::getenv("HSS_DE_SEED");
    seed = s ? std::strtoull(s, nullptr, 10) : static_cast<uint64_t>(time(nullptr));
    midpointStream = 0;
    const Gather* g = Gather::getInstance();
    HSS_CHECK(hss_de_workspace_size(1, static_cast<int32_t>(g->getSamplesPerTrace()),
                                    traveltime->getModel(),
                                    static_cast<int32_t>(individualsPerPopulation), &workspaceBytes));
    if (!workspace) HIP_CHECK(hipMalloc(&workspace, workspaceBytes));
    if (!result) {
        HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&result),
                            sizeof(float) * traveltime->getNumberOfResults() * g->getSamplesPerTrace()));
    }
}

void HipDifferentialEvolutionAlgorithm::startAllPopulations() {
    // the start population is drawn by the first generation's launch, from the midpoint's
    // Philox stream
    nextGeneration = 0;
}

void HipDifferentialEvolutionAlgorithm::selectTracesToBeUsedForMidpoint(float m0) {
    hipStream_t stream = hipBackendStream(deviceContext);
    Gather* gather = Gather::getInstance();
    hss_oct_candidates oct{};
    if (traveltime->getModel() == OCT) {
        // the traces some individual of the start population reaches
        oct.mode = HSS_OCT_DE_START;
        oct.NP = static_cast<int32_t>(individualsPerPopulation);
        for (unsigned int p = 0; p < 2; p++) {
            oct.lower[p] = traveltime->getLowerBoundForParameter(p);
            oct.upper[p] = traveltime->getUpperBoundForParameter(p);
        }
        oct.seed = seed;
        oct.stream_base = midpointStream;
    }
    const vector<unsigned char> mask = batch.selectTraces(
        deviceFilteredTracesDataMap, traveltime->getModel(), m0, gather->getApm(),
        traveltime->getReferenceHalfoffset(), traveltime->getModel() == OCT ? &oct : nullptr, stream);
    copyOnlySelectedTracesToDevice(mask);
}

void HipDifferentialEvolutionAlgorithm::computeSemblanceAtGpuForMidpoint(float m0) {
    hipStream_t stream = hipBackendStream(deviceContext);
    Gather* gather = Gather::getInstance();
    const unsigned int gen = nextGeneration++;
    const unsigned int np = traveltime->getNumberOfParameters();
    if (!filteredTracesCount) {
        if (gen + 1 == generations) {
            HIP_CHECK(hipMemsetAsync(result, 0, sizeof(float) * (2 + np) * gather->getSamplesPerTrace(),
                                     stream));
            HIP_CHECK(hipStreamSynchronize(stream));
        }
        return;
    }
    if (gen == 0) batch.reset(stream);
    float lower[3] = {0.0f, 0.0f, 0.0f}, upper[3] = {0.0f, 0.0f, 0.0f};
    for (unsigned int p = 0; p < np; p++) {
        lower[p] = traveltime->getLowerBoundForParameter(p);
        upper[p] = traveltime->getUpperBoundForParameter(p);
    }
    const hss_traces traces = batch.traces(deviceFilteredTracesDataMap, filteredTracesCount);
    const hss_midpoints mids = batch.midpoint(m0, filteredTracesCount,
                                              traveltime->getReferenceHalfoffset(),
                                              gather->getApm(), stream);
    HSS_CHECK(hss_de_semblance(&traces, &mids, traveltime->getModel(), lower, upper,
                               static_cast<int32_t>(individualsPerPopulation),
                               static_cast<int32_t>(generations), static_cast<int32_t>(gen),
                               static_cast<int32_t>(gen + 1), seed, midpointStream,
                               hipBackendVariant(), workspace, workspaceBytes, result, nullptr,
                               nullptr, batch.notUsed(), stream));
    HIP_CHECK(hipStreamSynchronize(stream));
}

// mutation, crossover and selection run inside each generation's launch
void HipDifferentialEvolutionAlgorithm::mutateAllPopulations() {}
void HipDifferentialEvolutionAlgorithm::crossoverPopulationIndividuals() {}
void HipDifferentialEvolutionAlgorithm::advanceGeneration() {}

void HipDifferentialEvolutionAlgorithm::selectBestIndividuals(vector<float>& resultArrays) {
    hipStream_t stream = hipBackendStream(deviceContext);
    const size_t n = traveltime->getNumberOfResults() * Gather::getInstance()->getSamplesPerTrace();
    if (resultArrays.size() != n) throw invalid_argument("result array of the wrong size");
    HIP_CHECK(hipMemcpy(resultArrays.data(), result, n * sizeof(float), hipMemcpyDeviceToHost));
    batch.storeNotUsed(deviceNotUsedCountArray.get(), stream);
    midpointStream++;
}
