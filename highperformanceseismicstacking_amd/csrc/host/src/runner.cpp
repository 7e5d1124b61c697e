// Single-host driver (see runner.hpp).
#include "runner.hpp"

#include <cerrno>
#include <chrono>
#include <iostream>
#include <sstream>
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
        LOGI("worker: set-up " << setUp.count() << " s, waiting on the shared queue and results "
                               << blocked.count() << " s");
    } catch (...) {
        std::lock_guard<std::mutex> lock(ref->resultSetMutex);
        if (!ref->workerError) ref->workerError = std::current_exception();
    }
}

namespace {

// Wall-clock time of the named phases of one run, logged together at the end.
class PhaseTimes {
  public:
    template <class F>
    void run(const char *name, F &&f) {
        const auto t = std::chrono::steady_clock::now();
        f();
        phases.emplace_back(name, std::chrono::steady_clock::now() - t);
    }
    void log() const {
        std::ostringstream s;
        s << "phases (s):";
        for (const auto &p : phases) s << ' ' << p.first << ' ' << p.second.count() << ';';
        LOGI(s.str());
    }

  private:
    std::vector<std::pair<const char *, Seconds>> phases;
};

}  // namespace

void SingleHostRunner::loadLine() {
    parser->readGather();
    Gather *gather = Gather::getInstance();
    // every CDP of the line, in ascending midpoint order, is one unit of work
    for (const auto &kv : gather->getCdps()) midpointQueue.push(kv.first);
    resultSet = std::make_unique<ResultSet>(traveltime->getNumberOfResults(), gather->getSamplesPerTrace());
}

void SingleHostRunner::runWorkers(unsigned int devices) {
    std::vector<std::thread> pool;
    pool.reserve(devices);
    for (unsigned int d = 0; d < devices; d++) pool.emplace_back(workerThread, this);
    for (auto &t : pool) t.join();
    if (workerError) std::rethrow_exception(workerError);
}

void SingleHostRunner::writeResults(Dumper &dumper) {
    dumpResultSet(dumper, parser->getInputFilePath(), *traveltime, *resultSet);
    std::ostringstream means;
    means << "statistics (mean over midpoints):";
    for (unsigned int i = 0; i < kStatisticCount; i++) {
        const StatisticResult r = static_cast<StatisticResult>(i);
        means << ' ' << statisticName(r) << ' ' << resultSet->get(r).getAverageOfAllMidpoints() << ';';
    }
    LOGI(means.str());
}

// The exit status of a run that ended with the exception being handled, as the
// reference's programs return it (SingleHostRunner.cpp:183-195): a usage error prints the
// message and the help and exits 0, a runtime error (no device, unreadable file) -ENODEV,
// any other failure -1.
int SingleHostRunner::exitStatusOfFailure() const {
    try {
        throw;
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
    } catch (...) {
        std::cout << "unknown failure" << std::endl;
        return -1;
    }
}

int SingleHostRunner::main(int argc, const char *argv[]) {
    const auto begin = std::chrono::steady_clock::now();
    try {
        PhaseTimes phases;
        unsigned int devices = 0;
        phases.run("devices", [&] { devices = getNumOfDevices(); });
        phases.run("arguments", [&] {
            parser->parseArguments(argc, argv);
            traveltime.reset(parser->parseTraveltime());
        });
        Dumper dumper(parser->getOutputDirectory(), parser->getFilename(), parser->getParserType(),
                      traveltime->getTraveltimeWord());
        dumper.createDir();
        phases.run("read", [&] { loadLine(); });
        phases.run("search", [&] { runWorkers(devices); });
        phases.run("write", [&] { writeResults(dumper); });
        phases.log();
        const Seconds total = std::chrono::steady_clock::now() - begin;
        LOGI("Results written to " << dumper.getOutputDirectoryPath());
        LOGI("run finished in " << total.count() << " s on " << devices << " device(s), "
                                << parser->getGpuThreadCount() << " threads per block requested");
        return 0;
    } catch (...) {
        return exitStatusOfFailure();
    }
}

HipSingleHostRunner::HipSingleHostRunner(Parser *parser)
    : SingleHostRunner(parser, HipComputeAlgorithmBuilder::getInstance(),
                       HipDeviceContextBuilder::getInstance()) {}

unsigned int HipSingleHostRunner::getNumOfDevices() const {
    int count = 0;
    const int rc = hss_device_count(&count);
    if (rc != HSS_OK || count == 0) {
        throw std::runtime_error("no HIP device is visible to this process");
    }
    return static_cast<unsigned int>(count);
}
