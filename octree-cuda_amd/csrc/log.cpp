#include "log.h"

#include <atomic>
#include <cstdio>

namespace oc {

static std::atomic<int> gLevel{LOG_WARN};

void setLogLevel(int level)
{
    gLevel = level;
}

void log(int level, const std::string& message)
{
    if (level < gLevel.load())
        return;
    static const char* names[] = {"debug", "info", "warning", "error"};
    std::fprintf(stderr, "[octree-amd] [%s] %s\n", names[level < 0 ? 0 : (level > 3 ? 3 : level)], message.c_str());
}

}  // namespace oc
