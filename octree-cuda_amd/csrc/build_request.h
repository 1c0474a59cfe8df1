// BuildRequest: what a build service needs from a request (a configured session): its
// configuration, its cloud, a place for the finished result and the request bookkeeping. Session
// implements it; the service's CPU test harness has a stand-in.
#pragma once

#include <atomic>
#include <chrono>
#include <cstdint>
#include <memory>
#include <string>

#include "build_result.h"

namespace oc {

class BuildRequest
{
public:
    virtual ~BuildRequest() = default;

    virtual void validate() const                   = 0;  // throws when the configuration is invalid
    virtual const CloudInfo& cloudInfo() const      = 0;
    virtual const ProcessingInfo& processingInfo() const = 0;
    virtual const uint8_t* hostCloud() const        = 0;
    virtual const uint8_t* deviceCloud() const      = 0;
    virtual void adoptResult(std::unique_ptr<HostBuildResult> result) = 0;

    enum class Request
    {
        NONE,
        QUEUED,
        RUNNING,
        DONE,
        FAILED
    };
    // Written by the service under its lock. state and owner are also read by the request's
    // destructor, which holds no lock, so they are atomic; the rest is read after wait() returns.
    struct RequestInfo
    {
        std::atomic<Request> state{Request::NONE};
        std::string error;
        uint64_t footprint = 0;
        std::atomic<class ServiceBase*> owner{nullptr};  // the service the request was submitted to
        std::chrono::steady_clock::time_point submitted, started, finished;
    };
    RequestInfo& request() { return itsRequest; }

protected:
    RequestInfo itsRequest;
};

}  // namespace oc
