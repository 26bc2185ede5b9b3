// Error translation at the C ABI: every entry point body runs inside
// INSPIRE_TRY, which turns an inspire::Error into its status code and keeps
// the message for inspire_last_error().
#pragma once
#include <new>
#include <stdexcept>

#include "context.hpp"

#define INSPIRE_TRY(...)                                    \
    try {                                                   \
        __VA_ARGS__                                         \
    } catch (const inspire::Error& e) {                     \
        inspire::set_last_error(e.msg);                     \
        return e.code;                                      \
    } catch (const std::bad_alloc&) {                       \
        inspire::set_last_error("host allocation failed");  \
        return INSPIRE_ERR_OUT_OF_MEMORY;                   \
    } catch (const std::exception& e) {                     \
        inspire::set_last_error(e.what());                  \
        return INSPIRE_ERR_DEVICE;                          \
    }
