// Per-midpoint tasks (see task.hpp).
#include "task.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "hip_backend.hpp"
#include "logger.hpp"

namespace {

Parser *parserOfProgram(int program) {
    switch (program) {
        case 0: return LinearSearchParser::getInstance();
        case 1: return DifferentialEvolutionParser::getInstance();
        case 2: return StretchFreeParser::getInstance();
        default: throw std::invalid_argument("program must be 0 (linear search), 1 (DE) or 2 (stretch-free)");
    }
}

float readFloat(const uint8_t *p) {
    float v;
    std::memcpy(&v, p, sizeof v);
    return v;
}

}  // namespace

TaskJob::TaskJob(int program, int argc, const char **argv) : parser(parserOfProgram(program)) {
    parser->parseArguments(argc, argv);
    traveltime.reset(parser->parseTraveltime());
    Gather::resetInstance();
    parser->readGather();
    const Gather *gather = Gather::getInstance();
    samples = gather->getSamplesPerTrace();
    for (const auto &kv : gather->getCdps()) midpoints.push_back(kv.first);
    resultSet = std::make_unique<ResultSet>(traveltime->getNumberOfResults(), samples);
    dumper = std::make_unique<Dumper>(parser->getOutputDirectory(), parser->getFilename(),
                                      parser->getParserType(), traveltime->getTraveltimeWord());
    dumper->createDir();
    LOGI("task job: " << midpoints.size() << " midpoints, records of " << resultBytes() << " bytes");
}

size_t TaskJob::resultBytes() const {
    return sizeof(float) * (1 + static_cast<size_t>(traveltime->getNumberOfResults()) * samples +
                            kStatisticCount);
}

bool TaskJob::nextTask(std::vector<uint8_t> &task) {
    std::lock_guard<std::mutex> lock(mu);
    if (next == midpoints.size()) return false;
    task.resize(sizeof(float));
    std::memcpy(task.data(), &midpoints[next++], sizeof(float));
    return true;
}

void TaskJob::commitTask(const uint8_t *record, size_t length) {
    if (!record || length != resultBytes())
        throw std::invalid_argument("a result record of " + std::to_string(resultBytes()) +
                                    " bytes was expected, got " + std::to_string(length));
    const float m0 = readFloat(record);
    if (!std::binary_search(midpoints.begin(), midpoints.end(), m0))
        throw std::invalid_argument("the result record's midpoint is not a CDP of this line");
    const size_t nvalues = static_cast<size_t>(traveltime->getNumberOfResults()) * samples;
    std::vector<float> values(nvalues);
    std::memcpy(values.data(), record + sizeof(float), nvalues * sizeof(float));
    std::lock_guard<std::mutex> lock(mu);
    resultSet->setAllResultsForMidpoint(m0, values);
    const uint8_t *stats = record + sizeof(float) * (1 + nvalues);
    for (unsigned s = 0; s < kStatisticCount; s++)
        resultSet->setStatisticalResultForMidpoint(m0, static_cast<StatisticResult>(s),
                                                   readFloat(stats + s * sizeof(float)));
    committed++;
}

std::string TaskJob::commitJob() {
    std::lock_guard<std::mutex> lock(mu);
    dumpResultSet(*dumper, parser->getInputFilePath(), *traveltime, *resultSet);
    LOGI("task job: " << committed << " result records written to " << dumper->getOutputDirectoryPath());
    return dumper->getOutputDirectoryPath();
}

TaskWorker::TaskWorker(TaskJob &j, unsigned int device) : job(j) {
    std::shared_ptr<DeviceContext> context(HipDeviceContextBuilder::getInstance()->build(device));
    algorithm.reset(job.getParser()->parseComputeAlgorithm(HipComputeAlgorithmBuilder::getInstance(),
                                                           context, job.getTraveltime()));
}

void TaskWorker::run(const uint8_t *task, size_t length, std::vector<uint8_t> &result) {
    if (!task || length != job.taskBytes())
        throw std::invalid_argument("a task record holds one float midpoint");
    if (!algorithm->isSetUp()) algorithm->setUp();
    const float m0 = readFloat(task);
    std::vector<MidpointOutput> out;
    algorithm->computeSemblanceAndParametersForMidpoints({m0}, out);
    if (out.size() != 1) throw std::logic_error("the backend returned no result for the task");
    const MidpointOutput &o = out[0];
    if (sizeof(float) * (1 + o.results.size() + kStatisticCount) != job.resultBytes())
        throw std::logic_error("the backend's result arrays do not match the traveltime's");
    result.resize(job.resultBytes());
    uint8_t *p = result.data();
    std::memcpy(p, &o.m0, sizeof(float));
    p += sizeof(float);
    std::memcpy(p, o.results.data(), o.results.size() * sizeof(float));
    p += o.results.size() * sizeof(float);
    std::memcpy(p, o.statistics.data(), kStatisticCount * sizeof(float));
}
