// Results, statistics and output files (see output.hpp).
#include "output.hpp"

#include <ctime>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "logger.hpp"

const std::string &statisticName(StatisticResult r) {
    static const std::string names[] = {"efficiency", "interpolations_per_sec", "selected_traces",
                                        "total_selection_kernel_exec_time",
                                        "total_kernel_exec_time"};
    return names[static_cast<int>(r)];
}

const std::vector<float> &MidpointResult::get(float m0) const {
    auto it = values.find(m0);
    if (it == values.end()) throw std::invalid_argument("no result stored for this midpoint");
    return it->second;
}

void MidpointResult::save(float m0, std::vector<float>::const_iterator begin,
                          std::vector<float>::const_iterator end) {
    values[m0].assign(begin, end);
}

float StatisticalMidpointResult::getStatisticalResultForMidpoint(float m0) const {
    auto it = values.find(m0);
    if (it == values.end()) throw std::invalid_argument("no result stored for this midpoint");
    return it->second;
}

float StatisticalMidpointResult::getAverageOfAllMidpoints() const {
    if (values.empty()) return 0;
    float sum = 0;
    for (const auto &kv : values) {
        if (kv.second != 0) sum += kv.second;
    }
    return sum / static_cast<float>(values.size());
}

const StatisticalMidpointResult &ResultSet::get(StatisticResult r) const {
    auto it = statistics.find(static_cast<int>(r));
    if (it == statistics.end()) throw std::invalid_argument("no midpoint has a value of this statistic");
    return it->second;
}

void ResultSet::setAllResultsForMidpoint(float m0, const std::vector<float> &array) {
    if (array.size() < resultArray.size() * static_cast<size_t>(samplesPerResult)) {
        throw std::invalid_argument("Result array shorter than results x samples.");
    }
    for (size_t i = 0; i < resultArray.size(); i++) {
        auto begin = array.cbegin() + static_cast<std::ptrdiff_t>(i * samplesPerResult);
        resultArray[i].save(m0, begin, begin + samplesPerResult);
    }
}

// ---------------------------------------------------------------------------------
Dumper::Dumper(const std::string &path, const std::string &dataFile,
               const std::string &computeMethod, const std::string &traveltime) {
    std::time_t t = std::time(nullptr);
    std::tm now{};
    localtime_r(&t, &now);
    std::ostringstream s;
    s << path << "/" << dataFile << "_" << computeMethod << "_" << traveltime << "_"
      << std::put_time(&now, "%Y%m%d_%H%M") << "/";
    outputDirectoryPath = s.str();
}

void Dumper::createDir() const {
    LOGD("output directory: " << outputDirectoryPath);
    if (std::filesystem::exists(outputDirectoryPath)) return;
    if (!std::filesystem::create_directory(outputDirectoryPath)) {
        throw std::runtime_error("cannot create the output directory " + outputDirectoryPath);
    }
}

static std::ofstream open_or_throw(const std::string &file, const std::string &what) {
    std::ofstream out(file, std::ios::binary);
    if (!out.is_open()) throw std::runtime_error("cannot write " + what + " to " + file);
    return out;
}

void Dumper::dumpGatherParameters(const std::string &file) const {
    auto out = open_or_throw(outputDirectoryPath + "gather.txt", "gather parameters");
    out << "==== Gather Summary ====" << std::endl;
    out << "Input file = " << file << std::endl;
    out << Gather::getInstance()->toString() << std::endl;
}

void Dumper::dumpAlgorithm(const std::string &summary) const {
    auto out = open_or_throw(outputDirectoryPath + "algorithm.txt", "algorithm parameters");
    out << "==== Compute Algorithm Summary ====" << std::endl;
    out << summary << std::endl;
}

void Dumper::dumpTraveltime(const Traveltime *model) const {
    auto out = open_or_throw(outputDirectoryPath + "traveltime.txt", "traveltime model parameters");
    out << "==== Model Summary ====" << std::endl;
    out << model->toString() << std::endl;
}

void Dumper::dumpResult(const std::string &resultName, const MidpointResult &result) const {
    auto out = open_or_throw(outputDirectoryPath + resultName + ".su", "result data");
    for (const auto &kv : Gather::getInstance()->getCdps()) {
        const std::vector<float> &samples = result.get(kv.first);
        const trace_info_t info = kv.second.getCdpInfo();
        out.write(reinterpret_cast<const char *>(&info), sizeof(trace_info_t));
        out.write(reinterpret_cast<const char *>(samples.data()),
                  static_cast<std::streamsize>(samples.size() * sizeof(float)));
    }
}

void Dumper::dumpStatisticalResult(const std::string &name, const StatisticalMidpointResult &r) const {
    auto out = open_or_throw(outputDirectoryPath + name + ".csv", "statistical result data");
    for (const auto &kv : Gather::getInstance()->getCdps()) {
        out << kv.first << "," << r.getStatisticalResultForMidpoint(kv.first) << std::endl;
    }
}

void dumpResultSet(const Dumper &dumper, const std::string &inputPath, const Traveltime &model,
                   const ResultSet &results) {
    dumper.dumpGatherParameters(inputPath);
    dumper.dumpTraveltime(&model);
    for (unsigned int r = 0; r < model.getNumberOfResults(); r++)
        dumper.dumpResult(model.getDescriptionForResult(r), results.getArrayForResult(r));
    for (unsigned int i = 0; i < static_cast<unsigned int>(StatisticResult::CNT); i++) {
        const StatisticResult r = static_cast<StatisticResult>(i);
        dumper.dumpStatisticalResult(statisticName(r), results.get(r));
    }
}
