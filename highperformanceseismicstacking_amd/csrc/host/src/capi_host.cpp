// C API of the host layer (include/hss_host.h).
#include "hss_host.h"
#include "hip_backend.hpp"
#include "semblance_op.hpp"
#include "task.hpp"

#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "logger.hpp"
#include "model.hpp"
#include "output.hpp"
#include "parser.hpp"
#include "traveltime.hpp"

static thread_local std::string last_error;

template <class F>
static int guarded(F &&f) {
    try {
        f();
        last_error.clear();
        return 0;
    } catch (const std::invalid_argument &e) {
        last_error = e.what();
        return -1;
    } catch (const std::logic_error &e) {
        last_error = e.what();
        return -2;
    } catch (const std::runtime_error &e) {
        last_error = e.what();
        return -3;
    } catch (const std::exception &e) {
        last_error = e.what();
        return -4;
    }
}

// A path into a caller's buffer, NUL-terminated and truncated to its capacity.
static void copy_path(const std::string &d, char *dir, int32_t dir_capacity) {
    if (!dir || dir_capacity <= 0) return;
    const size_t n = std::min(d.size(), static_cast<size_t>(dir_capacity - 1));
    std::memcpy(dir, d.data(), n);
    dir[n] = 0;
}

extern "C" {

const char *hss_host_last_error(void) { return last_error.c_str(); }

int hss_host_read_gather(const char *path, float aph, float apm, float tau, float azimuth_deg,
                         const float *lowest, const float *highest, int32_t *ns, float *dt,
                         int32_t *tau_index, int32_t *w, int32_t *ntr, int32_t *ncdp) {
    return guarded([&] {
        Logger::getInstance()->setVerbosity(LogLevel::ERROR);
        Gather::resetInstance();
        Gather *g = Gather::getInstance();
        g->setAph(aph);
        g->setApm(apm);
        g->setTau(tau);
        g->setAzimuthInDegree(azimuth_deg);
        if (lowest) g->setLowestAllowedMidpoint(*lowest);
        if (highest) g->setBiggestAllowedMidpoint(*highest);
        g->readGatherFromFile(path);
        *ns = static_cast<int32_t>(g->getSamplesPerTrace());
        *dt = g->getSamplePeriodInSeconds();
        *tau_index = g->getTauIndexDisplacement();
        *w = static_cast<int32_t>(g->getWindowSize());
        *ntr = static_cast<int32_t>(g->getTotalTracesCount());
        *ncdp = static_cast<int32_t>(g->getTotalCdpsCount());
    });
}

int hss_host_gather_arrays(float *mid, float *hoff, float *samples, float *cdp_m0, int32_t *cdp_traces) {
    return guarded([&] {
        const Gather *g = Gather::getInstance();
        const unsigned ns = g->getSamplesPerTrace();
        for (unsigned i = 0; i < g->getTotalTracesCount(); i++) {
            const Trace &t = g->getTraceAtIndex(i);
            if (mid) mid[i] = t.getMidpoint();
            if (hoff) hoff[i] = t.getHalfoffset();
            if (samples) std::memcpy(samples + static_cast<size_t>(i) * ns, t.getSamples().data(), ns * sizeof(float));
        }
        size_t k = 0;
        for (const auto &kv : g->getCdps()) {
            if (cdp_m0) cdp_m0[k] = kv.first;
            if (cdp_traces) cdp_traces[k] = static_cast<int32_t>(kv.second.getTraceCount());
            k++;
        }
    });
}

int hss_host_traveltime_bounds(int model, const float *lower, const float *upper, int32_t n,
                               float v0, float ain, float bpctg, float *out_lower, float *out_upper) {
    return guarded([&] {
        std::vector<float> lo(lower, lower + n), hi(upper, upper + n);
        TraveltimeBuilder *b = TraveltimeBuilder::getInstance();
        std::unique_ptr<Traveltime> t;
        if (model == CMP) t.reset(b->buildCommonMidPoint(lo, hi));
        else if (model == ZOCRS) t.reset(b->buildZeroOffsetCommonReflectionSurface(lo, hi, v0, ain, bpctg));
        else if (model == OCT) t.reset(b->buildOffsetContinuationTrajectory(lo, hi, 0.0f));
        else throw std::invalid_argument("Invalid traveltime model");
        for (unsigned p = 0; p < t->getNumberOfParameters(); p++) {
            out_lower[p] = t->getLowerBoundForParameter(p);
            out_upper[p] = t->getUpperBoundForParameter(p);
        }
    });
}

int hss_host_xcd_balanced_order(const int32_t *work, int32_t n, int32_t *order) {
    return guarded([&] {
        if (n < 0) throw std::invalid_argument("n must be >= 0");
        if (n > 0 && (!work || !order)) throw std::invalid_argument("work and order are required");
        const std::vector<int32_t> w(work, work + n);
        const std::vector<int32_t> o = xcd_balanced_order(w);
        std::copy(o.begin(), o.end(), order);
    });
}

int hss_host_dump_results(const char *out_path, const char *input_path, const char *data_name,
                          const char *method, int model, const float *lower, const float *upper,
                          int32_t nbounds, float v0, float ain, float bpctg, float h0,
                          const float *values, const float *statistics, char *dir,
                          int32_t dir_capacity) {
    return guarded([&] {
        if (!out_path || !input_path || !data_name || !method || !values || !statistics)
            throw std::invalid_argument("null argument");
        if (nbounds < 0 || (nbounds > 0 && (!lower || !upper)))
            throw std::invalid_argument("bounds missing");
        const Gather *g = Gather::getInstance();
        std::vector<float> lo(lower, lower + nbounds), hi(upper, upper + nbounds);
        TraveltimeBuilder *b = TraveltimeBuilder::getInstance();
        std::unique_ptr<Traveltime> tt;
        if (model == CMP) tt.reset(b->buildCommonMidPoint(lo, hi));
        else if (model == ZOCRS) tt.reset(b->buildZeroOffsetCommonReflectionSurface(lo, hi, v0, ain, bpctg));
        else if (model == OCT) tt.reset(b->buildOffsetContinuationTrajectory(lo, hi, h0));
        else throw std::invalid_argument("Invalid traveltime model");
        const unsigned nres = tt->getNumberOfResults(), ns = g->getSamplesPerTrace();
        const unsigned nstat = static_cast<unsigned>(StatisticResult::CNT);
        ResultSet results(nres, ns);
        size_t k = 0;
        for (const auto &kv : g->getCdps()) {
            const float *v = values + k * nres * ns;
            results.setAllResultsForMidpoint(kv.first, std::vector<float>(v, v + nres * ns));
            for (unsigned s = 0; s < nstat; s++)
                results.setStatisticalResultForMidpoint(kv.first, static_cast<StatisticResult>(s),
                                                        statistics[k * nstat + s]);
            k++;
        }
        Dumper dumper(out_path, data_name, method, tt->getTraveltimeWord());
        dumper.createDir();
        dumpResultSet(dumper, input_path, *tt, results);
        copy_path(dumper.getOutputDirectoryPath(), dir, dir_capacity);
    });
}

int hss_host_parse_command_line(int program, int argc, const char **argv) {
    return guarded([&] {
        std::unique_ptr<Parser> p;
        if (program == 0) p.reset(new LinearSearchParser());
        else if (program == 1) p.reset(new DifferentialEvolutionParser());
        else if (program == 2) p.reset(new StretchFreeParser());
        else throw std::invalid_argument("unknown program");
        p->parseArguments(argc, argv);
        std::unique_ptr<Traveltime> t(p->parseTraveltime());
        Logger::getInstance()->setVerbosity(LogLevel::ERROR);
    });
}


// ---- per-midpoint tasks ----------------------------------------------------------------
struct hss_host_job {
    std::unique_ptr<TaskJob> job;
};
struct hss_host_worker {
    std::unique_ptr<TaskWorker> worker;
};

int hss_host_job_open(int program, int argc, const char **argv, hss_host_job **job) {
    return guarded([&] {
        if (!job) throw std::invalid_argument("null job handle");
        *job = nullptr;
        auto j = std::make_unique<hss_host_job>();
        j->job = std::make_unique<TaskJob>(program, argc, argv);
        *job = j.release();
    });
}

int hss_host_job_record_sizes(const hss_host_job *job, int32_t *task_bytes, int32_t *result_bytes,
                              int32_t *tasks) {
    return guarded([&] {
        if (!job || !task_bytes || !result_bytes || !tasks) throw std::invalid_argument("null argument");
        *task_bytes = static_cast<int32_t>(job->job->taskBytes());
        *result_bytes = static_cast<int32_t>(job->job->resultBytes());
        *tasks = static_cast<int32_t>(job->job->taskCount());
    });
}

int hss_host_job_next_task(hss_host_job *job, void *task, int32_t capacity, int32_t *length) {
    return guarded([&] {
        if (!job || !task || !length) throw std::invalid_argument("null argument");
        if (capacity < static_cast<int32_t>(job->job->taskBytes()))
            throw std::invalid_argument("task buffer too small");
        std::vector<uint8_t> t;
        *length = 0;
        if (!job->job->nextTask(t)) return;
        std::memcpy(task, t.data(), t.size());
        *length = static_cast<int32_t>(t.size());
    });
}

int hss_host_job_commit_task(hss_host_job *job, const void *result, int32_t length) {
    return guarded([&] {
        if (!job || length < 0) throw std::invalid_argument("null job or negative length");
        job->job->commitTask(static_cast<const uint8_t *>(result), static_cast<size_t>(length));
    });
}

int hss_host_job_commit(hss_host_job *job, char *dir, int32_t dir_capacity) {
    return guarded([&] {
        if (!job) throw std::invalid_argument("null job");
        copy_path(job->job->commitJob(), dir, dir_capacity);
    });
}

int hss_host_job_close(hss_host_job *job) {
    return guarded([&] { delete job; });
}

int hss_host_worker_open(hss_host_job *job, int32_t device, hss_host_worker **worker) {
    return guarded([&] {
        if (!job || !worker || device < 0) throw std::invalid_argument("null argument or negative device");
        *worker = nullptr;
        auto w = std::make_unique<hss_host_worker>();
        w->worker = std::make_unique<TaskWorker>(*job->job, static_cast<unsigned>(device));
        *worker = w.release();
    });
}

int hss_host_worker_run(hss_host_worker *worker, const void *task, int32_t length, void *result,
                        int32_t capacity, int32_t *result_length) {
    return guarded([&] {
        if (!worker || !result || !result_length || length < 0)
            throw std::invalid_argument("null argument or negative length");
        std::vector<uint8_t> r;
        worker->worker->run(static_cast<const uint8_t *>(task), static_cast<size_t>(length), r);
        if (capacity < static_cast<int32_t>(r.size())) throw std::invalid_argument("result buffer too small");
        std::memcpy(result, r.data(), r.size());
        *result_length = static_cast<int32_t>(r.size());
    });
}

int hss_host_worker_close(hss_host_worker *worker) {
    return guarded([&] { delete worker; });
}

int hss_host_semblance_op(const hss_traces *traces, const hss_midpoints *mids, int model,
                          const float *d_points, const float *d_gradS, float *d_S,
                          float *d_gradPoints, void *stream) {
    return guarded([&] {
        if (!traces || !mids) throw std::invalid_argument("null traces or midpoints");
        SemblanceOp op(*traces, *mids, model, static_cast<hipStream_t>(stream));
        op.forward(d_points, d_S);
        op.backward(d_gradS, d_gradPoints);
        HIP_ASSERT(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    });
}

}  // extern "C"
