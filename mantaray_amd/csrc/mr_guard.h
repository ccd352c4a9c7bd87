// mr_guard.h -- no C++ exception leaves an extern "C" entry point.
//
// Every entry point of the C ABI wraps its body in MR_GUARD_BEGIN /
// MR_GUARD_END: an exception thrown inside (std::bad_alloc from a host
// allocation sized by a caller or a file header, std::length_error, ...)
// becomes a status code and a message in mr_last_error(), instead of
// unwinding through the C boundary and terminating the process.
#pragma once

#include <new>
#include <stdexcept>
#include <string>

#include "mantaray_capi.h"

namespace mr {

int set_error(int code, const std::string &msg);   // mr_capi.cpp

// the status for the exception being handled (call inside a catch block)
inline int status_of_current_exception() noexcept
{
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return set_error(MR_ERR_OUT_OF_MEMORY, "out of host memory");
    } catch (const std::length_error &e) {
        return set_error(MR_ERR_OUT_OF_MEMORY, std::string("allocation too large: ") + e.what());
    } catch (const std::exception &e) {
        return set_error(MR_ERR_INVALID_ARGUMENT, std::string("unexpected error: ") + e.what());
    } catch (...) {
        return set_error(MR_ERR_INVALID_ARGUMENT, "unexpected error");
    }
}

}  // namespace mr

#define MR_GUARD_BEGIN try {
#define MR_GUARD_END                                \
    }                                               \
    catch (...)                                     \
    {                                               \
        return mr::status_of_current_exception();   \
    }
