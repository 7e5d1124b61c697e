// Traveltime models: parameter names, bounds and result naming of the reference's
// Traveltime hierarchy (common/include/traveltime/*.hpp), re-implemented.
//
// Results of a model are (coherence, stack, parameters...), the parameters being
// CMP {velocity}; ZO-CRS {velocity, a, b}; OCT {velocity, slope}.  ZO-CRS takes only the
// velocity bounds from the command line; A and B are bounded by +-2 sin(ain)/v0 and
// +-bpctg*4/vmin^2 (TraveltimeBuilder.cpp:45-76).  The stretch-free wrapper reports one
// result parameter, "n", while keeping the wrapped model's parameter count.
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "hss.h"

// traveltime_t (common/include/capability.h:3-7)
enum traveltime_t { CMP = HSS_CMP, ZOCRS = HSS_ZOCRS, OCT = HSS_OCT };

enum class SemblanceCommonResult { SEMBL, STACK, CNT };

class TraveltimeParameter {
  public:
    explicit TraveltimeParameter(const std::string &description) : description(description) {}
    void updateMinimum(float v) { bounds.first = v; }
    void updateMaximum(float v) { bounds.second = v; }
    const std::string &getParameterDescription() const { return description; }
    float getMinimum() const { return bounds.first; }
    float getMaximum() const { return bounds.second; }

  private:
    std::string description;
    std::pair<float, float> bounds{0.0f, 0.0f};
};

class Traveltime {
  public:
    virtual ~Traveltime() = default;

    unsigned int getIndexForCommonResult(SemblanceCommonResult r) const {
        return static_cast<unsigned int>(r);
    }
    virtual unsigned int getNumberOfParameters() const {
        return static_cast<unsigned int>(parameters.size());
    }
    unsigned int getNumberOfCommonResults() const {
        return static_cast<unsigned int>(SemblanceCommonResult::CNT);
    }
    virtual unsigned int getNumberOfResults() const {
        return getNumberOfCommonResults() + getNumberOfParameters();
    }
    const std::string getDescriptionForParameter(unsigned int i) const;
    const std::string getDescriptionForResult(unsigned int i) const;
    virtual float getReferenceHalfoffset() const { return referenceHalfOffset; }
    void updateLowerBoundForParameter(unsigned int p, float v) { parameters.at(p).updateMinimum(v); }
    void updateUpperBoundForParameter(unsigned int p, float v) { parameters.at(p).updateMaximum(v); }
    float getLowerBoundForParameter(unsigned int p) const { return parameters.at(p).getMinimum(); }
    float getUpperBoundForParameter(unsigned int p) const { return parameters.at(p).getMaximum(); }

    virtual const std::string getTraveltimeWord() const = 0;
    virtual void updateReferenceHalfoffset(float h0) = 0;
    virtual traveltime_t getModel() const = 0;
    virtual const std::string toString() const = 0;

  protected:
    explicit Traveltime(std::vector<TraveltimeParameter> p) : parameters(std::move(p)) {}
    std::string boundsToString() const;
    std::vector<TraveltimeParameter> parameters;
    float referenceHalfOffset = 0.0f;
};

class CommonMidPoint : public Traveltime {
  public:
    static unsigned int VELOCITY;
    CommonMidPoint() : Traveltime({TraveltimeParameter("velocity")}) {}
    traveltime_t getModel() const override { return CMP; }
    const std::string getTraveltimeWord() const override { return "cmp"; }
    void updateReferenceHalfoffset(float) override {}
    const std::string toString() const override;
};

class CommonReflectionSurface : public Traveltime {
  public:
    static unsigned int VELOCITY, A, B;
    CommonReflectionSurface()
        : Traveltime({TraveltimeParameter("velocity"), TraveltimeParameter("a"),
                      TraveltimeParameter("b")}) {}
    float getAinInDegrees() const { return ainInDegrees; }
    float getAinInRad() const { return ainInRadian; }
    float getRatioForB() const { return bRatio; }
    float getReferenceVelocity() const { return referenceVelocity; }
    void setAinInDegrees(float deg);
    void setAinInRad(float rad);
    void setRatioForB(float p) { bRatio = p; }
    void setReferenceVelocity(float v) { referenceVelocity = v; }
    traveltime_t getModel() const override { return ZOCRS; }
    const std::string getTraveltimeWord() const override { return "zocrs"; }
    void updateReferenceHalfoffset(float) override {}
    const std::string toString() const override;

  private:
    float ainInDegrees = 0, ainInRadian = 0, referenceVelocity = 0, bRatio = 0;
};

class OffsetContinuationTrajectory : public Traveltime {
  public:
    static unsigned int VELOCITY, SLOPE;
    OffsetContinuationTrajectory()
        : Traveltime({TraveltimeParameter("velocity"), TraveltimeParameter("slope")}) {}
    traveltime_t getModel() const override { return OCT; }
    const std::string getTraveltimeWord() const override { return "oct"; }
    void updateReferenceHalfoffset(float h0) override { referenceHalfOffset = h0; }
    const std::string toString() const override;
};

// Stretch-free view of a model: results (coherence, stack, n), parameters those of the
// wrapped model (StretchFreeTraveltimeWrapper.cpp:3-29).
class StretchFreeTraveltimeWrapper : public Traveltime {
  public:
    explicit StretchFreeTraveltimeWrapper(Traveltime *wrapped)
        : Traveltime({TraveltimeParameter("n")}), wrapped(wrapped) {}
    unsigned int getNumberOfParameters() const override { return wrapped->getNumberOfParameters(); }
    // coherence, stack and n for every wrapped model (the reference sizes this as
    // 2 + parameters, which only fits CMP)
    unsigned int getNumberOfResults() const override { return getNumberOfCommonResults() + 1; }
    float getReferenceHalfoffset() const override { return wrapped->getReferenceHalfoffset(); }
    traveltime_t getModel() const override { return wrapped->getModel(); }
    const std::string getTraveltimeWord() const override {
        return "stretch_free_" + wrapped->getTraveltimeWord();
    }
    void updateReferenceHalfoffset(float h0) override { wrapped->updateReferenceHalfoffset(h0); }
    const std::string toString() const override { return wrapped->toString(); }
    // the wrapped model's parameter bounds (for the stretch-free input files' meaning)
    const Traveltime &getWrapped() const { return *wrapped; }

  private:
    std::unique_ptr<Traveltime> wrapped;
};

class TraveltimeBuilder {
  public:
    static TraveltimeBuilder *getInstance();
    Traveltime *buildCommonMidPoint(const std::vector<float> &lower, const std::vector<float> &upper);
    Traveltime *buildZeroOffsetCommonReflectionSurface(const std::vector<float> &lower,
                                                       const std::vector<float> &upper, float v0,
                                                       float aIn, float bPctg);
    Traveltime *buildOffsetContinuationTrajectory(const std::vector<float> &lower,
                                                  const std::vector<float> &upper, float h0);
};
