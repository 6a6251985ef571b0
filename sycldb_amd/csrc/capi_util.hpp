// Conversion of C++ errors to C-ABI status codes.
#pragma once

#include <new>
#include <stdexcept>
#include <string>

#include "sdb_common.hpp"

namespace sdb {

void set_last_error(const std::string &msg);

template <class F> int guarded(F &&f)
{
    try {
        f();
        set_last_error("");
        return SDB_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &e) {
        set_last_error(std::string("out of memory: ") + e.what());
        return SDB_ERR_OOM;
    } catch (const std::invalid_argument &e) {
        set_last_error(e.what());
        return SDB_ERR_INVALID_ARG;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return SDB_ERR_RUNTIME;
    }
}

inline void require(bool cond, const char *msg)
{
    if (!cond) fail(SDB_ERR_INVALID_ARG, msg);
}

inline hipStream_t as_stream(sdb_stream s) { return reinterpret_cast<hipStream_t>(s); }

} // namespace sdb
