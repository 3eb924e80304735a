// Error plumbing for the C ABI: a thread-local message and a guard that turns
// any C++ exception into a status code, so nothing unwinds across the ABI.
#pragma once

#include <exception>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/ommhip.h"

namespace ommhip {

void set_last_error(const std::string& message);

inline omm_status fail(omm_status code, const std::string& message) {
    set_last_error(message);
    return code;
}

// Thrown inside the library; carries the status to return.
struct Error : std::runtime_error {
    omm_status code;
    Error(omm_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <class F>
omm_status guarded(F&& body) {
    try {
        set_last_error("");
        return body();
    } catch (const Error& e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc&) {
        return fail(OMM_ERR_INVALID, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(OMM_ERR_INVALID, e.what());
    } catch (...) {
        return fail(OMM_ERR_INVALID, "unknown error");
    }
}

}  // namespace ommhip
