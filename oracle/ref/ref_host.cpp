/*
 * Reference-built checker, host side -- TEST INFRASTRUCTURE ONLY.
 *
 * Links the reference's SU reader and geometry (common/src/model/{Gather,Trace,Cdp,Utils}.cpp)
 * and traveltime builder (common/src/traveltime/*.cpp), compiled from /root/reference
 * unmodified, behind a small command-line driver written here:
 *
 *   ref_host gather <in.su> <aph> <apm> <tau> <azimuth> <lowest|-> <highest|-> <out-prefix>
 *       reads the file with Gather::readGatherFromFile and writes <out-prefix>.txt
 *       (scalars, the CDP map and every kept trace's midpoint/half offset as hex floats)
 *       and <out-prefix>.bin (the kept traces' samples, float32, in gather order);
 *   ref_host dump <in.su> <aph> <apm> <tau> <azimuth> <cmp|zocrs|oct> <nb> <lower x nb>
 *                 <upper x nb> <v0> <ain> <bpctg> <h0> <out-path> <data-name> <method>
 *                 <values.bin> <stats.bin>
 *       reads the file, builds the traveltime with TraveltimeBuilder, fills a ResultSet
 *       with the given values (float32 [cdp][result][ns], CDPs in map order) and
 *       statistics (float32 [cdp][5]) and writes them with the reference's own Dumper
 *       (common/src/output/Dumper.cpp, ResultSet.cpp and the result classes, unmodified)
 *       in the order SingleHostRunner::main writes them; prints the output directory;
 *   ref_host zocrs <vmin> <vmax> <v0> <ain> <bpctg>
 *   ref_host oct <vmin> <vmax> <smin> <smax> <h0>
 *       print the parameter bounds TraveltimeBuilder derives, as hex floats.
 *
 * Gather is a process-wide singleton, so each read runs in a process of its own.
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "common/include/model/Gather.hpp"
#include "common/include/output/Dumper.hpp"
#include "common/include/output/Logger.hpp"
#include "common/include/semblance/result/ResultSet.hpp"
#include "common/include/semblance/result/StatisticResult.hpp"
#include "common/include/traveltime/TraveltimeBuilder.hpp"

static int do_gather(int argc, char **argv)
{
    if (argc != 10) return 2;
    Logger::getInstance()->setVerbosity(LogLevel::ERROR);
    Gather *g = Gather::getInstance();
    g->setAph(strtof(argv[3], nullptr));
    g->setApm(strtof(argv[4], nullptr));
    g->setTau(strtof(argv[5], nullptr));
    g->setAzimuthInDegree(strtof(argv[6], nullptr));
    if (strcmp(argv[7], "-") != 0) g->setLowestAllowedMidpoint(strtof(argv[7], nullptr));
    if (strcmp(argv[8], "-") != 0) g->setBiggestAllowedMidpoint(strtof(argv[8], nullptr));
    g->readGatherFromFile(argv[2]);
    std::string prefix = argv[9];
    FILE *t = fopen((prefix + ".txt").c_str(), "w");
    std::ofstream b(prefix + ".bin", std::ios::binary);
    fprintf(t, "ns %u\ndt %a\ntau %d\nw %u\nntr %u\nncdp %u\n", g->getSamplesPerTrace(),
            g->getSamplePeriodInSeconds(), g->getTauIndexDisplacement(), g->getWindowSize(),
            g->getTotalTracesCount(), g->getTotalCdpsCount());
    for (const auto &it : g->getCdps()) {
        fprintf(t, "cdp %a\n", it.first);
    }
    for (unsigned int i = 0; i < g->getTotalTracesCount(); i++) {
        const Trace &tr = g->getTraceAtIndex(i);
        fprintf(t, "trace %a %a\n", tr.getMidpoint(), tr.getHalfoffset());
        b.write(reinterpret_cast<const char *>(tr.getSamples().data()),
                tr.getSamples().size() * sizeof(float));
    }
    fclose(t);
    return 0;
}

static std::vector<float> read_floats(const char *path)
{
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    const std::streamsize bytes = in.tellg();
    in.seekg(0);
    std::vector<float> v(static_cast<size_t>(bytes) / sizeof(float));
    in.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(float)));
    return v;
}

static int do_dump(int argc, char **argv)
{
    if (argc < 9) return 2;
    Logger::getInstance()->setVerbosity(LogLevel::ERROR);
    Gather *g = Gather::getInstance();
    g->setAph(strtof(argv[3], nullptr));
    g->setApm(strtof(argv[4], nullptr));
    g->setTau(strtof(argv[5], nullptr));
    g->setAzimuthInDegree(strtof(argv[6], nullptr));
    const std::string model = argv[7];
    const int nb = atoi(argv[8]);
    if (argc != 9 + 2 * nb + 9) return 2;
    std::vector<float> lo, hi;
    for (int i = 0; i < nb; i++) lo.push_back(strtof(argv[9 + i], nullptr));
    for (int i = 0; i < nb; i++) hi.push_back(strtof(argv[9 + nb + i], nullptr));
    char **rest = argv + 9 + 2 * nb;
    const float v0 = strtof(rest[0], nullptr), ain = strtof(rest[1], nullptr);
    const float bpctg = strtof(rest[2], nullptr), h0 = strtof(rest[3], nullptr);
    g->readGatherFromFile(argv[2]);
    TraveltimeBuilder *b = TraveltimeBuilder::getInstance();
    std::unique_ptr<Traveltime> tt;
    if (model == "cmp") tt.reset(b->buildCommonMidPoint(lo, hi));
    else if (model == "zocrs") tt.reset(b->buildZeroOffsetCommonReflectionSurface(lo, hi, v0, ain, bpctg));
    else if (model == "oct") tt.reset(b->buildOffsetContinuationTrajectory(lo, hi, h0));
    else return 2;
    const unsigned nres = tt->getNumberOfResults(), ns = g->getSamplesPerTrace();
    const std::vector<float> values = read_floats(rest[7]), stats = read_floats(rest[8]);
    const unsigned nstat = static_cast<unsigned>(StatisticResult::CNT);
    if (values.size() != g->getCdps().size() * nres * ns || stats.size() != g->getCdps().size() * nstat)
        return 3;
    ResultSet results(nres, ns);
    size_t k = 0;
    for (const auto &it : g->getCdps()) {
        std::vector<float> row(values.begin() + k * nres * ns, values.begin() + (k + 1) * nres * ns);
        results.setAllResultsForMidpoint(it.first, row);
        for (unsigned s = 0; s < nstat; s++)
            results.setStatisticalResultForMidpoint(it.first, static_cast<StatisticResult>(s),
                                                    stats[k * nstat + s]);
        k++;
    }
    Dumper dumper(rest[4], rest[5], rest[6], tt->getTraveltimeWord());
    dumper.createDir();
    dumper.dumpGatherParameters(argv[2]);
    dumper.dumpTraveltime(tt.get());
    for (unsigned i = 0; i < nres; i++)
        dumper.dumpResult(tt->getDescriptionForResult(i), results.getArrayForResult(i));
    for (unsigned s = 0; s < nstat; s++) {
        const StatisticResult r = static_cast<StatisticResult>(s);
        dumper.dumpStatisticalResult(STATISTIC_NAME_MAP[r], results.get(r));
    }
    printf("%s\n", dumper.getOutputDirectoryPath().c_str());
    return 0;
}

static int do_traveltime(int argc, char **argv)
{
    std::unique_ptr<Traveltime> tt;
    if (!strcmp(argv[1], "zocrs") && argc == 7) {
        std::vector<float> lo = {strtof(argv[2], nullptr)}, hi = {strtof(argv[3], nullptr)};
        tt.reset(TraveltimeBuilder::getInstance()->buildZeroOffsetCommonReflectionSurface(
            lo, hi, strtof(argv[4], nullptr), strtof(argv[5], nullptr), strtof(argv[6], nullptr)));
    } else if (!strcmp(argv[1], "oct") && argc == 7) {
        std::vector<float> lo = {strtof(argv[2], nullptr), strtof(argv[4], nullptr)};
        std::vector<float> hi = {strtof(argv[3], nullptr), strtof(argv[5], nullptr)};
        tt.reset(TraveltimeBuilder::getInstance()->buildOffsetContinuationTrajectory(
            lo, hi, strtof(argv[6], nullptr)));
    } else {
        return 2;
    }
    for (unsigned int p = 0; p < tt->getNumberOfParameters(); p++) {
        printf("%s %a %a\n", tt->getDescriptionForParameter(p).c_str(),
               tt->getLowerBoundForParameter(p), tt->getUpperBoundForParameter(p));
    }
    printf("h0 %a\n", tt->getReferenceHalfoffset());
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    if (!strcmp(argv[1], "gather")) return do_gather(argc, argv);
    if (!strcmp(argv[1], "dump")) return do_dump(argc, argv);
    return do_traveltime(argc, argv);
}
