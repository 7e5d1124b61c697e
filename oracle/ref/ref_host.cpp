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
#include "common/include/output/Logger.hpp"
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
    return do_traveltime(argc, argv);
}
