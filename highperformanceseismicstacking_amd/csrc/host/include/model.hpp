// Prestack data model and SU input: the reference's Trace / Cdp / Gather / Utils
// interface (common/include/model/*.hpp), re-implemented.
//
// Behaviour kept from the reference because results depend on it:
//  * SU traces are raw native-endian records: a 240-byte header, then ns float32
//    samples; no byte swap, no 3600-byte SEG-Y file header (Trace.cpp:45-61).
//  * Coordinates: f = scalco >= 0 ? scalco : -1/scalco; m = f*(gx+sx)*0.5 and
//    h = f*(gx-sx)*0.5 per axis, projected as x*sin(az) + y*cos(az) (Utils.cpp:8-52).
//  * Reading (Gather.cpp:49-126): the first trace fixes ns and dt; a CDP entry is
//    created for every midpoint seen, before the filters; traces with a different ns,
//    |h| > aph or a midpoint outside [lowest, highest] are dropped; the default lowest
//    midpoint is FLT_MIN, so midpoints <= 0 are dropped; traces end sorted by (m, h).
//  * The reference's read loop makes one extra read at end of file, which inserts a CDP
//    at m = 0 unless one exists.  That CDP is kept here too (it has no traces and is
//    processed and written like the others).  Its header is never read in the
//    reference (a default-constructed Trace leaves it uninitialised), and the reference
//    built with g++ leaves the previous record's header bytes in it; the reader here
//    gives it exactly those bytes (the last record's header, then offset = 0 and the
//    source/receiver coordinates averaged as for every CDP), so the written files match
//    the reference's byte for byte.
#pragma once

#include <cstdint>
#include <fstream>
#include <map>
#include <memory>
#include <string>
#include <vector>

// The 240-byte SU trace header (field layout of the SU/SEG-Y trace header).
#pragma pack(push, 1)
struct trace_info_t {
    int32_t tracl, tracr, fldr, tracf, ep, cdp, cdpt;
    int16_t trid, nvs, nhs, duse;
    int32_t offset, gelev, selev, sdepth, gdel, sdel, swdep, gwdep;
    int16_t scalel, scalco;
    int32_t sx, sy, gx, gy;
    int16_t counit, wevel, swevel, sut, gut, sstat, gstat, tstat, laga, lagb, delrt, muts, mute;
    uint16_t ns, dt;
    int16_t gain, igc, igi, corr, sfs, sfe, slen, styp, stas, stae, tatyp, afilf, afils,
        nofilf, nofils, lcf, hcf, lcs, hcs, year, day, hour, minute, sec, timbas, trwf, grnors,
        grnofr, grnlof, gaps, otrav;
    float d1, f1, d2, f2, ungpow, unscale;
    int32_t ntr;
    int16_t mark, shortpad;
    int16_t unass[14];
};
#pragma pack(pop)
static_assert(sizeof(trace_info_t) == 240, "SU trace header must be 240 bytes");

constexpr float MICRO2SECONDS = 1e-6f;

enum class Axis { X, Y };

namespace Utils {
float halfoffset_axis(const trace_info_t &info, Axis axis);
float midpoint_axis(const trace_info_t &info, Axis axis);
float halfoffset(const trace_info_t &info, float azimuthInRad);
float midpoint(const trace_info_t &info, float azimuthInRad);
float degreesToRad(float degree);
float radToDegrees(float rad);
}  // namespace Utils

class Trace {
  public:
    Trace() = default;
    Trace(const std::vector<float> &samples, const trace_info_t &info, float azimuthInRad);

    float getHalfoffset() const { return halfoffset; }
    float getMidpoint() const { return midpoint; }
    const trace_info_t &getInfo() const { return info; }
    unsigned int getSampleCount() const { return static_cast<unsigned int>(samples.size()); }
    const std::vector<float> &getSamples() const { return samples; }
    bool isMidpointInRange(float lower, float upper) const;

    // Reads one record.  Returns false (leaving the trace without samples, at m = h = 0)
    // when the stream holds no further header; its header then holds `previous` with
    // whatever bytes a partial header read overwrote.  Throws std::runtime_error when the
    // samples are truncated.
    bool read(std::ifstream &in, float azimuthInRad, const trace_info_t &previous);

    bool operator<(const Trace &other) const;

  private:
    float halfoffset = 0.0f, midpoint = 0.0f;
    std::vector<float> samples;
    trace_info_t info{};
};

class Cdp {
  public:
    Cdp() = default;
    Cdp(float midpoint, const trace_info_t &info) : midpoint(midpoint), cdpInfo(info) {}
    const trace_info_t &getCdpInfo() const { return cdpInfo; }
    float getMidpoint() const { return midpoint; }
    unsigned int getTraceCount() const { return traceCount; }
    unsigned int getSampleCount() const { return sampleCount; }
    void incrementTraceCountBy(unsigned int n) { traceCount += n; }
    void incrementSampleCountBy(unsigned int n) { sampleCount += n; }

  private:
    float midpoint = 0.0f;
    trace_info_t cdpInfo{};
    unsigned int sampleCount = 0, traceCount = 0;
};

class Gather {
  public:
    static Gather *getInstance();
    // Drops the process-wide instance (tests and repeated runs in one process).
    static void resetInstance();

    void readGatherFromFile(const std::string &path);
    void add(const Trace &trace);

    float getApm() const { return apm; }
    float getAph() const { return aph; }
    float getAzimuthInDegrees() const { return azimuthInDegrees; }
    float getAzimuthInRad() const { return azimuthInRad; }
    float getTau() const { return tau; }
    int getTauIndexDisplacement() const { return tauIndexDisplacement; }
    unsigned int getWindowSize() const { return static_cast<unsigned int>(windowSize); }
    unsigned int getSamplesPerTrace() const { return samplesPerTrace; }
    unsigned short getSamplePeriod() const { return dt; }
    float getSamplePeriodInSeconds() const { return dtInSeconds; }
    unsigned int getTotalSamplesCount() const { return samplesCount; }
    unsigned int getTotalTracesCount() const { return static_cast<unsigned int>(traces.size()); }
    unsigned int getTotalCdpsCount() const { return static_cast<unsigned int>(cdps.size()); }
    float getLowestAllowedMidpoint() const { return lowestAllowedMidpoint; }
    float getBiggestAllowedMidpoint() const { return biggestAllowedMidpoint; }
    bool isGatherRead() const { return isRead; }
    const std::map<float, Cdp> &getCdps() const { return cdps; }
    const Trace &getTraceAtIndex(unsigned int i) const { return traces[i]; }
    float getMidpointOfTrace(unsigned int i) const { return traces[i].getMidpoint(); }
    float getHalfoffsetOfTrace(unsigned int i) const { return traces[i].getHalfoffset(); }

    void setApm(float v) { apm = v; }
    void setAph(float v) { aph = v; }
    void setTau(float v) { tau = v; }
    void setAzimuthInDegree(float deg);
    void setAzimuthInRad(float rad);
    void setLowestAllowedMidpoint(float v) { lowestAllowedMidpoint = v; }
    void setBiggestAllowedMidpoint(float v) { biggestAllowedMidpoint = v; }

    const std::string toString() const;

  private:
    float apm = 0, aph = 0, azimuthInDegrees = 0, azimuthInRad = 0;
    unsigned short dt = 0;
    float dtInSeconds = 0;
    float lowestAllowedMidpoint, biggestAllowedMidpoint;
    unsigned int samplesCount = 0, samplesPerTrace = 0;
    float tau = 0;
    int tauIndexDisplacement = 0, windowSize = 0;
    std::map<float, Cdp> cdps;
    std::vector<Trace> traces;
    bool isRead = false;

    Gather();
    static std::unique_ptr<Gather> instance;
};
