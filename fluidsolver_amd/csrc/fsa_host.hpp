// fsa_host.hpp — host-side error plumbing shared by the library's C entry points.
#pragma once

#include <new>
#include <stdexcept>
#include <string>

#include "../../include/fluidsolver_amd.h"

namespace fsa {

// text of the calling thread's last error (fsa_last_error)
extern thread_local std::string g_error;

struct FsaError : std::runtime_error {
    int32_t code;
    FsaError(int32_t c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// runs f, turning every exception into a status code and the thread's error text
template <class F>
int32_t guarded(F&& f) {
    try {
        f();
        return FSA_OK;
    } catch (const FsaError& e) {
        g_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_error = "host allocation failed";
        return FSA_E_NOMEM;
    } catch (const std::exception& e) {
        g_error = e.what();
        return FSA_E_INVALID;
    } catch (...) {
        g_error = "unknown error";
        return FSA_E_INVALID;
    }
}

}  // namespace fsa
