// Mapping of the runtime's exceptions to the C ABI's status codes
// (include/bw_cloth.h).
#pragma once
#include <exception>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/bw_cloth.h"

namespace bw
{

// status code and message of the exception in flight (call inside catch)
inline int status_of_current_exception(std::string &msg)
{
    try
    {
        throw;
    }
    catch (const std::invalid_argument &e)
    {
        msg = e.what();
        return BW_ERR_ARG;
    }
    catch (const std::runtime_error &e)
    {
        msg = e.what();
        if (msg.rfind("HIP error", 0) == 0)
            return BW_ERR_DEVICE;
        if (msg.rfind("step:", 0) == 0)
            return BW_ERR_NUMERIC;
        if (msg.rfind("config", 0) == 0 || msg.rfind("batch:", 0) == 0)
            return BW_ERR_ARG;
        if (msg.rfind("capacity:", 0) == 0)
            return BW_ERR_CAPACITY;
        return BW_ERR_STATE;
    }
    catch (const std::bad_alloc &)
    {
        msg = "out of host memory";
        return BW_ERR_CAPACITY;
    }
    catch (const std::exception &e)
    {
        msg = std::string("unexpected exception: ") + e.what();
        return BW_ERR_STATE;
    }
    catch (...)
    {
        msg = "unknown error";
        return BW_ERR_STATE;
    }
}

} // namespace bw
