// SU reading, trace geometry and the CDP map (see model.hpp for the behaviour kept).
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <sstream>
#include <stdexcept>

#include "logger.hpp"

// ---------------------------------------------------------------------------------
Logger *Logger::getInstance() {
    static Logger logger;
    return &logger;
}

void Logger::print(LogLevel level, const std::string &message) {
    if (level > getInstance()->getVerbosity()) return;
    static const char *const tags[] = {"ERROR: ", " INFO: ", "  LOW: ", " HIGH: ", "DEBUG: "};
    std::cout << tags[static_cast<int>(level)] << message << std::endl;
}

// ---------------------------------------------------------------------------------
namespace Utils {

static float coordinate_scale(const trace_info_t &info) {
    const float s = static_cast<float>(info.scalco);
    return info.scalco >= 0 ? s : -1.0f / s;
}

float halfoffset_axis(const trace_info_t &info, Axis axis) {
    const int d = axis == Axis::X ? info.gx - info.sx : info.gy - info.sy;
    return coordinate_scale(info) * static_cast<float>(d) * 0.5f;
}

float midpoint_axis(const trace_info_t &info, Axis axis) {
    const int d = axis == Axis::X ? info.gx + info.sx : info.gy + info.sy;
    return coordinate_scale(info) * static_cast<float>(d) * 0.5f;
}

float halfoffset(const trace_info_t &info, float az) {
    return halfoffset_axis(info, Axis::X) * std::sin(az) + halfoffset_axis(info, Axis::Y) * std::cos(az);
}

float midpoint(const trace_info_t &info, float az) {
    return midpoint_axis(info, Axis::X) * std::sin(az) + midpoint_axis(info, Axis::Y) * std::cos(az);
}

float degreesToRad(float degree) { return degree * static_cast<float>(M_PI) / 180.0f; }
float radToDegrees(float rad) { return rad * 180.0f / static_cast<float>(M_PI); }

}  // namespace Utils

// ---------------------------------------------------------------------------------
Trace::Trace(const std::vector<float> &s, const trace_info_t &i, float az)
    : halfoffset(Utils::halfoffset(i, az)), midpoint(Utils::midpoint(i, az)), samples(s), info(i) {}

bool Trace::isMidpointInRange(float lower, float upper) const {
    return midpoint >= lower && midpoint <= upper;
}

bool Trace::read(std::ifstream &in, float az, const trace_info_t &previous) {
    info = previous;
    if (!in.read(reinterpret_cast<char *>(&info), sizeof(trace_info_t))) {
        return false;
    }
    midpoint = Utils::midpoint(info, az);
    halfoffset = Utils::halfoffset(info, az);
    samples.resize(info.ns);
    LOGD("trace payload: " << samples.size() * sizeof(float) << " bytes");
    if (!in.read(reinterpret_cast<char *>(samples.data()),
                 static_cast<std::streamsize>(samples.size() * sizeof(float)))) {
        throw std::runtime_error("SU file ends inside the samples of a trace");
    }
    return true;
}

bool Trace::operator<(const Trace &o) const {
    return midpoint < o.midpoint || (midpoint == o.midpoint && halfoffset < o.halfoffset);
}

// ---------------------------------------------------------------------------------
std::unique_ptr<Gather> Gather::instance;

Gather::Gather()
    : lowestAllowedMidpoint(std::numeric_limits<float>::min()),
      biggestAllowedMidpoint(std::numeric_limits<float>::max()) {}

Gather *Gather::getInstance() {
    if (!instance) instance.reset(new Gather());
    return instance.get();
}

void Gather::resetInstance() { instance.reset(); }

void Gather::setAzimuthInDegree(float deg) {
    azimuthInDegrees = deg;
    azimuthInRad = Utils::degreesToRad(deg);
}

void Gather::setAzimuthInRad(float rad) {
    azimuthInRad = rad;
    azimuthInDegrees = Utils::radToDegrees(rad);
}

void Gather::add(const Trace &trace) {
    traces.push_back(trace);
    samplesCount += trace.getSampleCount();
}

void Gather::readGatherFromFile(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) {
        throw std::invalid_argument("cannot open input file " + path);
    }
    samplesCount = 0;
    bool more = true;
    trace_info_t previous{};
    while (more) {
        Trace trace;
        more = trace.read(in, azimuthInRad, previous);
        previous = trace.getInfo();
        // The record after the last trace is an empty trace at m = 0 carrying the last
        // record's header: it still creates a CDP entry, as the reference's loop does
        // (see model.hpp).
        const float m = trace.getMidpoint();
        const float h = trace.getHalfoffset();
        if (!samplesPerTrace) {
            samplesPerTrace = trace.getSampleCount();
            dt = trace.getInfo().dt;
            dtInSeconds = static_cast<float>(dt) * MICRO2SECONDS;
            tauIndexDisplacement = static_cast<int>(tau / dtInSeconds);
            windowSize = 2 * tauIndexDisplacement + 1;
        }
        if (cdps.find(m) == cdps.end()) {
            trace_info_t cdpInfo = trace.getInfo();
            cdpInfo.offset = 0;
            cdpInfo.sx = cdpInfo.gx = (trace.getInfo().sx + trace.getInfo().gx) >> 1;
            cdpInfo.sy = cdpInfo.gy = (trace.getInfo().sy + trace.getInfo().gy) >> 1;
            LOGH("new CDP at m = " << m);
            cdps[m] = Cdp(m, cdpInfo);
        }
        if (samplesPerTrace != trace.getSampleCount()) continue;
        if (std::abs(h) > aph) {
            LOGH("trace at m = " << m << " dropped: |h| = " << std::abs(h) << " beyond aph = " << aph);
            continue;
        }
        if (!trace.isMidpointInRange(lowestAllowedMidpoint, biggestAllowedMidpoint)) continue;
        add(trace);
        cdps[m].incrementTraceCountBy(1);
        cdps[m].incrementSampleCountBy(trace.getSampleCount());
    }
    std::sort(traces.begin(), traces.end());
    LOGI("gather: " << cdps.size() << " CDPs, " << traces.size() << " traces of "
                    << samplesPerTrace << " samples");
    isRead = true;
}

const std::string Gather::toString() const {
    std::ostringstream s;
    s << "Samples per trace = " << samplesPerTrace << std::endl;
    s << "Total traces count = " << getTotalTracesCount() << std::endl;
    s << "Total samples = " << samplesCount << std::endl;
    s << "CDPs = " << getTotalCdpsCount() << std::endl;
    s << "apm = " << apm << " m" << std::endl;
    s << "aph = " << aph << " m" << std::endl;
    s << "tau = " << tau << std::endl;
    s << "azimuth = " << azimuthInRad << " rad" << std::endl;
    if (lowestAllowedMidpoint > std::numeric_limits<float>::min())
        s << "Smallest allowed midpoint = " << lowestAllowedMidpoint << " m" << std::endl;
    if (biggestAllowedMidpoint < std::numeric_limits<float>::max())
        s << "Biggest allowed midpoint = " << biggestAllowedMidpoint << " m" << std::endl;
    return s.str();
}
