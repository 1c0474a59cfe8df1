// Minimal leveled logging to stderr (the reference uses spdlog with a global level,
// octreeApi.cpp:23-42).
#pragma once

#include <string>

namespace oc {

enum LogLevel
{
    LOG_DEBUG = 0,
    LOG_INFO  = 1,
    LOG_WARN  = 2,
    LOG_ERROR = 3
};

void setLogLevel(int level);
void log(int level, const std::string& message);

}  // namespace oc
