#include "hip/include/semblance/algorithm/HipComputeAlgorithmBuilder.hpp"
#include "hip/include/semblance/algorithm/HipDifferentialEvolutionAlgorithm.hpp"
#include "hip/include/semblance/algorithm/HipLinearSearchAlgorithm.hpp"
#include "hip/include/semblance/algorithm/HipStretchFreeAlgorithm.hpp"
#include "hip/include/semblance/data/HipDataContainerBuilder.hpp"

#include <sstream>
#include <stdexcept>

ComputeAlgorithmBuilder* HipComputeAlgorithmBuilder::getInstance() {
    static HipComputeAlgorithmBuilder builder;
    return &builder;
}

// The count checks of CudaComputeAlgorithmBuilder (logic_error on a mismatch).
static void expectOnePerParameter(const Traveltime& t, size_t given, const char* what) {
    if (given != t.getNumberOfParameters()) {
        ostringstream s;
        s << "the " << t.getTraveltimeWord() << " traveltime takes " << t.getNumberOfParameters()
          << " " << what << ", got " << given;
        throw logic_error(s.str());
    }
}

LinearSearchAlgorithm* HipComputeAlgorithmBuilder::buildLinearSearchAlgorithm(
    shared_ptr<Traveltime> traveltime, shared_ptr<DeviceContext> context, unsigned int threadCount,
    const vector<int>& discretizationArray) {
    expectOnePerParameter(*traveltime, discretizationArray.size(), "granularities");
    auto* algorithm = new HipLinearSearchAlgorithm(traveltime, context,
                                                   HipDataContainerBuilder::getInstance(), threadCount);
    for (unsigned int p = 0; p < traveltime->getNumberOfParameters(); p++) {
        algorithm->setDiscretizationGranularityForParameter(p, static_cast<unsigned int>(discretizationArray[p]));
    }
    return algorithm;
}

DifferentialEvolutionAlgorithm* HipComputeAlgorithmBuilder::buildDifferentialEvolutionAlgorithm(
    shared_ptr<Traveltime> traveltime, shared_ptr<DeviceContext> context, unsigned int threadCount,
    unsigned int generation, unsigned int individualsPerPopulation) {
    return new HipDifferentialEvolutionAlgorithm(traveltime, context, HipDataContainerBuilder::getInstance(),
                                                 threadCount, generation, individualsPerPopulation);
}

StretchFreeAlgorithm* HipComputeAlgorithmBuilder::buildStretchFreeAlgorithm(
    shared_ptr<Traveltime> traveltime, shared_ptr<DeviceContext> context, unsigned int threadCount,
    const vector<string>& parameterFileArray) {
    expectOnePerParameter(*traveltime, parameterFileArray.size(), "parameter files");
    return new HipStretchFreeAlgorithm(traveltime, context, HipDataContainerBuilder::getInstance(),
                                       threadCount, parameterFileArray);
}
