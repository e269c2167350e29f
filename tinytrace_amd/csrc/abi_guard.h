// The C ABI promises that no C++ exception crosses it.  Every extern "C"
// function that can reach a throwing call (standard containers and strings,
// std::thread, new without nothrow) is written as a function-try-block whose
// handler returns tt::exception_status: host allocation and host thread
// failures map to TT_OUT_OF_MEMORY, anything else to TT_INTERNAL_ERROR, with
// the exception's message stored where the function keeps its errors.
// Host-only: no HIP.
#pragma once

#include <exception>
#include <new>
#include <string>
#include <system_error>

#include "../../include/tinytrace.h"

namespace tt {

// *dst = text, unless that fails too (then *dst keeps what it held).
inline void set_error_nothrow(std::string* dst, const char* text) noexcept {
    if (!dst) return;
    try {
        *dst = text;
    } catch (...) {
    }
}

// Status of the exception being handled; call only inside a catch handler.
inline tt_status exception_status(std::string* error) noexcept {
    try {
        throw;
    } catch (const std::bad_alloc&) {
        set_error_nothrow(error, "host memory allocation failed");
        return TT_OUT_OF_MEMORY;
    } catch (const std::length_error&) {
        set_error_nothrow(error, "host memory allocation failed (size too large)");
        return TT_OUT_OF_MEMORY;
    } catch (const std::system_error& e) {
        // std::thread and std::mutex report resource exhaustion this way.
        set_error_nothrow(error, e.what());
        return TT_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_error_nothrow(error, e.what());
        return TT_INTERNAL_ERROR;
    } catch (...) {
        set_error_nothrow(error, "unknown C++ exception inside the library");
        return TT_INTERNAL_ERROR;
    }
}

}  // namespace tt
