// Single-host driver (see runner.hpp).
#include "runner.hpp"

#include <cerrno>
#include <chrono>
#include <iostream>
#include <stdexcept>
#include <thread>
#include <vector>

#include "hip_backend.hpp"
#include "logger.hpp"

SingleHostRunner::SingleHostRunner(Parser *p, ComputeAlgorithmBuilder *ab, DeviceContextBuilder *db)
    : parser(p), algorithmBuilder(ab), deviceContextBuilder(db) {}

ComputeAlgorithm *SingleHostRunner::getComputeAlgorithm() {
    std::lock_guard<std::mutex> lock(deviceMutex);
    std::shared_ptr<DeviceContext> context(deviceContextBuilder->build(deviceIndex++));
    return parser->parseComputeAlgorithm(algorithmBuilder, context, traveltime);
}

void SingleHostRunner::workerThread(SingleHostRunner *ref) {
    try {
        std::unique_ptr<ComputeAlgorithm> algorithm(ref->getComputeAlgorithm());
        Seconds setUp = Seconds::zero(), blocked = Seconds::zero();
        MEASURE_EXEC_TIME(setUp, algorithm->setUp());
        const unsigned int batchSize = BackendOptions::get().batchSize;
        std::vector<MidpointOutput> outputs;
        while (true) {
            std::vector<float> m0s;
            auto t = std::chrono::steady_clock::now();
            {
                std::lock_guard<std::mutex> lock(ref->queueMutex);
                while (!ref->midpointQueue.empty() && (batchSize == 0 || m0s.size() < batchSize)) {
                    m0s.push_back(ref->midpointQueue.front());
                    ref->midpointQueue.pop();
                }
            }
            blocked += std::chrono::steady_clock::now() - t;
            if (m0s.empty()) break;
            algorithm->computeSemblanceAndParametersForMidpoints(m0s, outputs);
            t = std::chrono::steady_clock::now();
            {
                std::lock_guard<std::mutex> lock(ref->resultSetMutex);
                for (const MidpointOutput &o : outputs) {
                    ref->resultSet->setAllResultsForMidpoint(o.m0, o.results);
                    for (unsigned s = 0; s < kStatisticCount; s++) {
                        ref->resultSet->setStatisticalResultForMidpoint(o.m0, static_cast<StatisticResult>(s),
                                                                        o.statistics[s]);
                    }
                }
            }
            blocked += std::chrono::steady_clock::now() - t;
        }
        LOGI("Set up time is " << setUp.count() << " s.");
        LOGI("Blocked time is " << blocked.count() << " s.");
    } catch (...) {
        std::lock_guard<std::mutex> lock(ref->resultSetMutex);
        if (!ref->workerError) ref->workerError = std::current_exception();
    }
}

int SingleHostRunner::main(int argc, const char *argv[]) {
    try {
        const auto start = std::chrono::steady_clock::now();
        Gather *gather = Gather::getInstance();
        Seconds initDriver = Seconds::zero(), parseArgs = Seconds::zero(), readGather = Seconds::zero();
        unsigned int devices = 0;
        MEASURE_EXEC_TIME(initDriver, devices = getNumOfDevices());
        MEASURE_EXEC_TIME(parseArgs, parser->parseArguments(argc, argv));
        traveltime.reset(parser->parseTraveltime());
        Dumper dumper(parser->getOutputDirectory(), parser->getFilename(), parser->getParserType(),
                      traveltime->getTraveltimeWord());
        dumper.createDir();
        MEASURE_EXEC_TIME(readGather, parser->readGather());
        for (const auto &kv : gather->getCdps()) midpointQueue.push(kv.first);
        resultSet = std::make_unique<ResultSet>(traveltime->getNumberOfResults(), gather->getSamplesPerTrace());

        LOGI("Starting worker threads...");
        const auto workersStart = std::chrono::steady_clock::now();
        std::vector<std::thread> threads;
        for (unsigned int d = 0; d < devices; d++) threads.emplace_back(workerThread, this);
        for (auto &t : threads) t.join();
        if (workerError) std::rethrow_exception(workerError);
        const Seconds workers = std::chrono::steady_clock::now() - workersStart;

        const auto dumpStart = std::chrono::steady_clock::now();
        dumper.dumpGatherParameters(parser->getInputFilePath());
        dumper.dumpTraveltime(traveltime.get());
        for (unsigned int i = 0; i < traveltime->getNumberOfResults(); i++) {
            dumper.dumpResult(traveltime->getDescriptionForResult(i), resultSet->getArrayForResult(i));
        }
        for (unsigned int i = 0; i < kStatisticCount; i++) {
            const StatisticResult r = static_cast<StatisticResult>(i);
            const StatisticalMidpointResult &sr = resultSet->get(r);
            dumper.dumpStatisticalResult(statisticName(r), sr);
            LOGI("Average of " << statisticName(r) << " is " << sr.getAverageOfAllMidpoints());
        }
        const Seconds dump = std::chrono::steady_clock::now() - dumpStart;
        const Seconds total = std::chrono::steady_clock::now() - start;
        LOGI("Driver init. duration is " << initDriver.count() << " s");
        LOGI("Read gather duration is " << readGather.count() << " s");
        LOGI("Parse args. duration is " << parseArgs.count() << " s");
        LOGI("Worker threads duration is " << workers.count() << " s");
        LOGI("Dump results duration is " << dump.count() << " s");
        LOGI("Thread/block is " << parser->getGpuThreadCount());
        LOGI("Results written to " << dumper.getOutputDirectoryPath());
        LOGI("It took " << total.count() << " s to compute.");
        return 0;
    } catch (const std::invalid_argument &e) {
        std::cout << e.what() << std::endl;
        parser->printHelp();
        return 0;
    } catch (const std::runtime_error &e) {
        std::cout << e.what() << std::endl;
        return -ENODEV;
    } catch (const std::exception &e) {
        std::cout << e.what() << std::endl;
        return -1;
    }
}

HipSingleHostRunner::HipSingleHostRunner(Parser *parser)
    : SingleHostRunner(parser, HipComputeAlgorithmBuilder::getInstance(),
                       HipDeviceContextBuilder::getInstance()) {}

unsigned int HipSingleHostRunner::getNumOfDevices() const {
    int count = 0;
    const int rc = hss_device_count(&count);
    if (rc != HSS_OK || count == 0) {
        throw std::runtime_error("No HIP devices found on your system. Exiting.");
    }
    return static_cast<unsigned int>(count);
}
