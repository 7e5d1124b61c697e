// Levelled logging to stdout with the reference's levels and macro names
// (common/include/output/Logger.hpp:8-46): ERROR < INFO < LOW < HIGH < DEBUG, INFO by
// default, chosen with --verbose 0..4.
#pragma once

#include <iostream>
#include <sstream>
#include <string>

enum class LogLevel { ERROR = 0, INFO = 1, LOW = 2, HIGH = 3, DEBUG = 4 };

class Logger {
  public:
    static Logger *getInstance();
    LogLevel getVerbosity() const { return verbosity; }
    void setVerbosity(LogLevel v) { verbosity = v; }
    static void print(LogLevel level, const std::string &message);

  private:
    LogLevel verbosity = LogLevel::INFO;
};

#define HSS_LOG(level, message)                                        \
    do {                                                               \
        if ((level) <= Logger::getInstance()->getVerbosity()) {        \
            std::ostringstream hss_log_stream_;                        \
            hss_log_stream_ << message;                                \
            Logger::print((level), hss_log_stream_.str());             \
        }                                                              \
    } while (0)

#define LOGE(m) HSS_LOG(LogLevel::ERROR, m)
#define LOGI(m) HSS_LOG(LogLevel::INFO, m)
#define LOGL(m) HSS_LOG(LogLevel::LOW, m)
#define LOGH(m) HSS_LOG(LogLevel::HIGH, m)
#define LOGD(m) HSS_LOG(LogLevel::DEBUG, m)
