// Command-line parsing (see parser.hpp).
#include "parser.hpp"

#include <algorithm>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "hip_backend.hpp"
#include "model.hpp"

void OptionTable::add(const std::string &name, Kind kind, const std::string &description,
                      bool required, const std::string &defaultValue) {
    options[name] = Option{kind, required, defaultValue, description};
}

std::map<std::string, std::vector<std::string>> OptionTable::parse(int argc, const char *argv[]) const {
    std::map<std::string, std::vector<std::string>> values;
    int i = 1;
    while (i < argc) {
        std::string tok = argv[i];
        if (tok.size() < 3 || tok.compare(0, 2, "--") != 0) {
            throw std::runtime_error("unexpected token '" + tok + "'");
        }
        std::string name = tok.substr(2), inlineValue;
        const bool hasInline = name.find('=') != std::string::npos;
        if (hasInline) {
            inlineValue = name.substr(name.find('=') + 1);
            name = name.substr(0, name.find('='));
        }
        auto it = options.find(name);
        if (it == options.end()) throw std::runtime_error("unrecognised option '--" + name + "'");
        const Kind kind = it->second.kind;
        const bool multi = kind == FLOATS || kind == INTS || kind == STRINGS;
        if (!multi && values.count(name)) {
            throw std::runtime_error("option '--" + name + "' cannot be specified more than once");
        }
        std::vector<std::string> &v = values[name];
        i++;
        if (kind == FLAG) {
            if (hasInline) throw std::runtime_error("option '--" + name + "' does not take a value");
            continue;
        }
        if (hasInline) v.push_back(inlineValue);
        while (i < argc && std::string(argv[i]).compare(0, 2, "--") != 0) {
            if (!multi && !v.empty()) break;
            v.push_back(argv[i]);
            i++;
        }
        if (v.empty()) throw std::runtime_error("the required argument for option '--" + name + "' is missing");
        if (!multi && v.size() != 1) throw std::runtime_error("option '--" + name + "' takes one value");
    }
    for (const auto &kv : options) {
        if (!values.count(kv.first) && !kv.second.defaultValue.empty()) {
            values[kv.first] = {kv.second.defaultValue};
        }
    }
    for (const auto &kv : options) {
        if (kv.second.required && !values.count(kv.first)) {
            throw std::runtime_error("the option '--" + kv.first + "' is required but missing");
        }
    }
    // every value must convert to its option's type, as Boost checks while storing
    for (const auto &kv : values) {
        const Kind kind = options.at(kv.first).kind;
        for (const std::string &s : kv.second) {
            size_t pos = 0;
            bool good = true;
            try {
                if (kind == FLOAT || kind == FLOATS) {
                    (void)std::stof(s, &pos);
                } else if (kind == INTS) {
                    (void)std::stoi(s, &pos);
                } else if (kind == UINT) {
                    good = std::stol(s, &pos) >= 0;
                } else {
                    pos = s.size();
                }
            } catch (const std::exception &) {
                good = false;
            }
            if (!good || pos != s.size()) {
                throw std::runtime_error("the argument ('" + s + "') for option '--" + kv.first + "' is invalid");
            }
        }
    }
    return values;
}

void OptionTable::print(std::ostream &out) const {
    out << "Allowed options:" << std::endl;
    for (const auto &kv : options) {
        std::string arg;
        switch (kv.second.kind) {
            case FLOAT: arg = " arg"; break;
            case UINT: arg = " arg"; break;
            case STRING: arg = " arg"; break;
            case FLOATS: case INTS: case STRINGS: arg = " arg..."; break;
            case FLAG: break;
        }
        if (!kv.second.defaultValue.empty()) arg += " (=" + kv.second.defaultValue + ")";
        out << "  --" << std::left << std::setw(28) << (kv.first + arg) << kv.second.description
            << std::endl;
    }
}

// ---------------------------------------------------------------------------------
Parser::Parser() {
    using K = OptionTable::Kind;
    arguments.add("aph", K::FLOAT, "APH constant.", true);
    arguments.add("apm", K::FLOAT, "APM constant.", true);
    arguments.add("ain", K::FLOAT, "Angle in degrees to be used during computation of A for ZO-CRS.");
    arguments.add("azimuth", K::FLOAT, "Azimuth angle in degrees.", true);
    arguments.add("bpctg", K::FLOAT, "Percentage of C that will be used to compute B for ZO-CRS.");
    arguments.add("h0", K::FLOAT, "h0 constant.");
    arguments.add("help", K::FLAG, "Produce help message.");
    arguments.add("highest-midpoint", K::FLOAT, "Higher bound value for a CDP midpoint.");
    arguments.add("input", K::STRING, "*.su input data file.", true);
    arguments.add("lower-bounds", K::FLOATS, "Lower bound for traveltime parameters.", true);
    arguments.add("lowest-midpoint", K::FLOAT, "Lower bound value for a CDP midpoint.");
    arguments.add("kernel-path", K::STRING, "Accepted for compatibility; kernels are prebuilt.");
    arguments.add("output", K::STRING, "*.su output data prefix.", true);
    arguments.add("tau", K::FLOAT, "Semblance processing window given by w = 2 * tau + 1.");
    arguments.add("thread-count", K::UINT, "GPU thread count (accepted; the kernels choose their shapes).",
                  false, std::to_string(DEFAULT_THREAD_COUNT));
    arguments.add("traveltime", K::STRING, "Traveltime traveltime to be used.", true);
    arguments.add("upper-bounds", K::FLOATS, "Upper bounds for traveltime parameters.", true);
    arguments.add("v0", K::FLOAT, "Velocity used to compute ZO-CRS A parameter.");
    arguments.add("verbose", K::UINT, "Verbosity level. Must be in the [0..4] integer range.");
    arguments.add("variant", K::STRING, "Window arithmetic: exact (reference order), fast (FMA) or kahan (FMA, compensated sums).", false, "fast");
    arguments.add("storage", K::STRING, "Trace sample storage on the device: f32, f16 or bf16.", false, "f32");
    arguments.add("batch-size", K::UINT, "Midpoints per kernel launch (0 = all).", false, "0");
    arguments.add("seed", K::STRING, "Differential-evolution seed (default: the clock).");
    arguments.add("polish-steps", K::UINT,
                  "Gauss-Newton steps on the semblance gradient applied to every winner (0 = none).",
                  false, "0");
}

void Parser::parseArguments(int argc, const char *argv[]) {
    try {
        argumentMap = arguments.parse(argc, argv);
        Logger::getInstance()->setVerbosity(getLogVerbosity());
        BackendOptions &o = BackendOptions::get();
        const std::string variant = getString("variant");
        if (variant != "exact" && variant != "fast" && variant != "kahan")
            throw std::runtime_error("--variant must be exact, fast or kahan");
        o.variant = variant == "exact" ? HSS_EXACT : (variant == "kahan" ? HSS_KAHAN : HSS_FAST);
        const std::string storage = getString("storage");
        if (storage == "f32") o.storage = HSS_F32;
        else if (storage == "f16") o.storage = HSS_F16;
        else if (storage == "bf16") o.storage = HSS_BF16;
        else throw std::runtime_error("--storage must be f32, f16 or bf16");
        o.batchSize = getUInt("batch-size");
        o.polishSteps = getUInt("polish-steps");
        o.seedGiven = has("seed");
        if (o.seedGiven) o.seed = std::stoull(getString("seed"));
    } catch (const std::exception &ex) {
        std::ostringstream e;
        e << "Unable to parse input arguments: " << ex.what();
        throw std::invalid_argument(e.str());
    }
}

float Parser::getFloat(const std::string &name) const {
    const std::string &s = argumentMap.at(name).at(0);
    size_t pos = 0;
    const float v = std::stof(s, &pos);
    if (pos != s.size()) throw std::runtime_error("the argument ('" + s + "') for option '--" + name + "' is invalid");
    return v;
}

unsigned int Parser::getUInt(const std::string &name) const {
    const std::string &s = argumentMap.at(name).at(0);
    size_t pos = 0;
    const long v = std::stol(s, &pos);
    if (pos != s.size() || v < 0) throw std::runtime_error("the argument ('" + s + "') for option '--" + name + "' is invalid");
    return static_cast<unsigned int>(v);
}

std::string Parser::getString(const std::string &name) const { return argumentMap.at(name).at(0); }

std::vector<float> Parser::getFloats(const std::string &name) const {
    std::vector<float> out;
    for (const auto &s : argumentMap.at(name)) {
        size_t pos = 0;
        out.push_back(std::stof(s, &pos));
        if (pos != s.size()) throw std::invalid_argument("the argument ('" + s + "') for option '--" + name + "' is invalid");
    }
    return out;
}

std::vector<int> Parser::getInts(const std::string &name) const {
    std::vector<int> out;
    for (const auto &s : argumentMap.at(name)) {
        size_t pos = 0;
        out.push_back(std::stoi(s, &pos));
        if (pos != s.size()) throw std::invalid_argument("the argument ('" + s + "') for option '--" + name + "' is invalid");
    }
    return out;
}

std::vector<std::string> Parser::getStrings(const std::string &name) const { return argumentMap.at(name); }

const std::string Parser::getFilename() const {
    return std::filesystem::path(getInputFilePath()).stem().string();
}

unsigned int Parser::getGpuThreadCount() const { return getUInt("thread-count"); }

const std::string Parser::getInputFilePath() const {
    if (has("input")) return getString("input");
    throw std::invalid_argument("Input data file not provided.");
}

const std::string Parser::getOutputDirectory() const {
    if (has("output")) return getString("output");
    throw std::invalid_argument("Output path not provided.");
}

LogLevel Parser::getLogVerbosity() const {
    if (has("verbose")) {
        return static_cast<LogLevel>(std::min(getUInt("verbose"), static_cast<unsigned int>(LogLevel::DEBUG)));
    }
    return LogLevel::INFO;
}

traveltime_t Parser::traveltimeFromString(const std::string &model) {
    if (model == "cmp") return CMP;
    if (model == "zocrs") return ZOCRS;
    if (model == "oct") return OCT;
    throw std::invalid_argument("Invalid traveltime: " + model);
}

Traveltime *Parser::parseTraveltime() const {
    if (!has("traveltime")) throw std::invalid_argument("Traveltime traveltime not provided.");
    TraveltimeBuilder *builder = TraveltimeBuilder::getInstance();
    const traveltime_t model = traveltimeFromString(getString("traveltime"));
    const std::vector<float> lower = getFloats("lower-bounds"), upper = getFloats("upper-bounds");
    if (model == CMP) return builder->buildCommonMidPoint(lower, upper);
    if (model == ZOCRS) {
        if (!has("ain") || !has("v0") || !has("bpctg")) {
            throw std::invalid_argument("Not all arguments have been provided.");
        }
        return builder->buildZeroOffsetCommonReflectionSurface(lower, upper, getFloat("v0"),
                                                               getFloat("ain"), getFloat("bpctg"));
    }
    return builder->buildOffsetContinuationTrajectory(lower, upper, has("h0") ? getFloat("h0") : 0.0f);
}

void Parser::printHelp() const { arguments.print(std::cout); }

void Parser::readGather() const {
    if (!has("apm") || !has("aph") || !has("tau")) {
        throw std::invalid_argument("Not all arguments to read the gather have been provided.");
    }
    Gather *gather = Gather::getInstance();
    gather->setApm(getFloat("apm"));
    gather->setAph(getFloat("aph"));
    gather->setTau(getFloat("tau"));
    if (has("azimuth")) gather->setAzimuthInDegree(getFloat("azimuth"));
    if (has("lowest-midpoint")) gather->setLowestAllowedMidpoint(getFloat("lowest-midpoint"));
    if (has("highest-midpoint")) gather->setBiggestAllowedMidpoint(getFloat("highest-midpoint"));
    gather->readGatherFromFile(getInputFilePath());
}

// ---------------------------------------------------------------------------------
LinearSearchParser::LinearSearchParser() {
    arguments.add("granularity", OptionTable::INTS, "Discretization granularities.", true);
}

Parser *LinearSearchParser::getInstance() {
    static LinearSearchParser parser;
    return &parser;
}

ComputeAlgorithm *LinearSearchParser::parseComputeAlgorithm(ComputeAlgorithmBuilder *builder,
                                                            std::shared_ptr<DeviceContext> context,
                                                            std::shared_ptr<Traveltime> traveltime) const {
    std::vector<int> granularity;
    if (has("granularity")) granularity = getInts("granularity");
    LOGD("grid search over " << granularity.size() << " parameter axes");
    ComputeAlgorithm *a = builder->buildLinearSearchAlgorithm(traveltime, context, getGpuThreadCount(), granularity);
    if (has("kernel-path")) a->compileKernels(getString("kernel-path"));
    return a;
}

DifferentialEvolutionParser::DifferentialEvolutionParser() {
    arguments.add("generations", OptionTable::UINT, "Number of generations to be used in differential evolution.", true);
    arguments.add("population-size", OptionTable::UINT, "Differential evolution population size.", true);
}

Parser *DifferentialEvolutionParser::getInstance() {
    static DifferentialEvolutionParser parser;
    return &parser;
}

ComputeAlgorithm *DifferentialEvolutionParser::parseComputeAlgorithm(
    ComputeAlgorithmBuilder *builder, std::shared_ptr<DeviceContext> context,
    std::shared_ptr<Traveltime> traveltime) const {
    if (!has("generations") || !has("population-size")) {
        throw std::logic_error("differential evolution needs --generations and --population-size");
    }
    const unsigned int generations = getUInt("generations");
    const unsigned int individuals = getUInt("population-size");
    LOGD("DE: " << generations << " generations of " << individuals << " individuals");
    ComputeAlgorithm *a = builder->buildDifferentialEvolutionAlgorithm(traveltime, context, getGpuThreadCount(),
                                                                       generations, individuals);
    if (has("kernel-path")) a->compileKernels(getString("kernel-path"));
    return a;
}

StretchFreeParser::StretchFreeParser() {
    arguments.add("stack-datafiles", OptionTable::STRINGS, "Datafiles to be used for stretch-free stack generation.", true);
}

Parser *StretchFreeParser::getInstance() {
    static StretchFreeParser parser;
    return &parser;
}

Traveltime *StretchFreeParser::parseTraveltime() const {
    return new StretchFreeTraveltimeWrapper(Parser::parseTraveltime());
}

ComputeAlgorithm *StretchFreeParser::parseComputeAlgorithm(ComputeAlgorithmBuilder *builder,
                                                           std::shared_ptr<DeviceContext> context,
                                                           std::shared_ptr<Traveltime> traveltime) const {
    std::vector<std::string> files;
    if (has("stack-datafiles")) files = getStrings("stack-datafiles");
    ComputeAlgorithm *a = builder->buildStretchFreeAlgorithm(traveltime, context, getGpuThreadCount(), files);
    if (has("kernel-path")) a->compileKernels(getString("kernel-path"));
    return a;
}
