// Shared host-side helpers of librln_estep: status handling and the reference's rounding macros.
#pragma once

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/rln_estep.h"

namespace rln {

// Thread-local last error text returned by rln_last_error().
void set_error(const char *fmt, ...);
const char *last_error();

// src/macros.h:197,212,226
inline int ROUND(double x) { return (x > 0) ? (int)(x + 0.5) : (int)(x - 0.5); }
inline int CEIL(double x) { return (x == (int)x) ? (int)x : ((x > 0) ? (int)(x + 1) : (int)x); }
inline int FLOOR(double x) { return (x == (int)x) ? (int)x : ((x > 0) ? (int)x : (int)(x - 1)); }

constexpr double kPi = 3.14159265358979323846;

}  // namespace rln

#define RLN_FAIL(code, ...)              \
    do {                                 \
        ::rln::set_error(__VA_ARGS__);   \
        return (code);                   \
    } while (0)
