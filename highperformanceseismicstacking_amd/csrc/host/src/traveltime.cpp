// Traveltime models and their builder (see traveltime.hpp).
#include "traveltime.hpp"

#include <cmath>
#include <sstream>
#include <stdexcept>

#include "logger.hpp"
#include "model.hpp"

unsigned int CommonMidPoint::VELOCITY = 0;
unsigned int CommonReflectionSurface::VELOCITY = 0;
unsigned int CommonReflectionSurface::A = 1;
unsigned int CommonReflectionSurface::B = 2;
unsigned int OffsetContinuationTrajectory::VELOCITY = 0;
unsigned int OffsetContinuationTrajectory::SLOPE = 1;

const std::string Traveltime::getDescriptionForParameter(unsigned int i) const {
    return parameters.at(i).getParameterDescription();
}

const std::string Traveltime::getDescriptionForResult(unsigned int i) const {
    static const char *const common[] = {"coherence", "stack"};
    const unsigned int nc = getNumberOfCommonResults();
    if (i < nc) return common[i];
    return parameters.at(i - nc).getParameterDescription();
}

// "Name = <word>", "h0 = <h0>", then "<param> in [lo, hi]" per parameter, the layout of
// the reference's traveltime.txt (CommonMidPoint.cpp:25-38 and twins).
std::string Traveltime::boundsToString() const {
    std::ostringstream s;
    s << "Name = " << getTraveltimeWord() << std::endl;
    s << "h0 = " << getReferenceHalfoffset() << std::endl;
    for (unsigned int i = 0; i < getNumberOfParameters(); i++) {
        s << getDescriptionForParameter(i) << " in [" << getLowerBoundForParameter(i) << ", "
          << getUpperBoundForParameter(i) << "]";
    }
    return s.str();
}

const std::string CommonMidPoint::toString() const { return boundsToString(); }

const std::string OffsetContinuationTrajectory::toString() const { return boundsToString(); }

void CommonReflectionSurface::setAinInDegrees(float deg) {
    ainInDegrees = deg;
    ainInRadian = Utils::degreesToRad(deg);
}

void CommonReflectionSurface::setAinInRad(float rad) {
    ainInRadian = rad;
    ainInDegrees = Utils::radToDegrees(rad);
}

const std::string CommonReflectionSurface::toString() const {
    std::ostringstream s;
    s << boundsToString();
    s << "v0 = " << referenceVelocity << std::endl;
    s << "ain = " << ainInRadian << " rad" << std::endl;
    s << "bpctg = " << bRatio << std::endl;
    return s.str();
}

// ---------------------------------------------------------------------------------
TraveltimeBuilder *TraveltimeBuilder::getInstance() {
    static TraveltimeBuilder builder;
    return &builder;
}

static void set_bounds(Traveltime *t, const std::vector<float> &lower,
                       const std::vector<float> &upper, const char *name) {
    const unsigned int n = t->getNumberOfParameters();
    if (lower.size() != n || upper.size() != n) {
        std::ostringstream e;
        e << "the " << name << " traveltime takes " << n << " lower and " << n << " upper bounds";
        delete t;
        throw std::logic_error(e.str());
    }
    for (unsigned int i = 0; i < n; i++) {
        t->updateLowerBoundForParameter(i, lower[i]);
        t->updateUpperBoundForParameter(i, upper[i]);
    }
}

Traveltime *TraveltimeBuilder::buildCommonMidPoint(const std::vector<float> &lower,
                                                   const std::vector<float> &upper) {
    Traveltime *t = new CommonMidPoint();
    set_bounds(t, lower, upper, "CMP");
    return t;
}

Traveltime *TraveltimeBuilder::buildZeroOffsetCommonReflectionSurface(
    const std::vector<float> &lower, const std::vector<float> &upper, float v0, float aIn,
    float bPctg) {
    if (lower.empty() || upper.empty()) {
        throw std::logic_error("the ZOCRS traveltime needs a velocity lower and upper bound");
    }
    auto *t = new CommonReflectionSurface();
    t->updateLowerBoundForParameter(CommonReflectionSurface::VELOCITY, lower[0]);
    t->updateUpperBoundForParameter(CommonReflectionSurface::VELOCITY, upper[0]);
    t->setReferenceVelocity(v0);
    t->setAinInDegrees(aIn);
    t->setRatioForB(bPctg);
    const float vmin = t->getLowerBoundForParameter(CommonReflectionSurface::VELOCITY);
    const float aMax = 2.0f * std::sin(t->getAinInRad()) / v0;
    const float bMax = bPctg * 4 / (vmin * vmin);
    LOGD("ZO-CRS search box: |A| <= " << aMax << ", |B| <= " << bMax);
    t->updateLowerBoundForParameter(CommonReflectionSurface::A, -aMax);
    t->updateUpperBoundForParameter(CommonReflectionSurface::A, aMax);
    t->updateLowerBoundForParameter(CommonReflectionSurface::B, -bMax);
    t->updateUpperBoundForParameter(CommonReflectionSurface::B, bMax);
    return t;
}

Traveltime *TraveltimeBuilder::buildOffsetContinuationTrajectory(const std::vector<float> &lower,
                                                                 const std::vector<float> &upper,
                                                                 float h0) {
    Traveltime *t = new OffsetContinuationTrajectory();
    set_bounds(t, lower, upper, "OCT");
    t->updateReferenceHalfoffset(h0);
    return t;
}
