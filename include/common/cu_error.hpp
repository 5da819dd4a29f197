// Fail-fast status check of the kept host API (common/cu_error.hpp:5-11):
// print "<name>: <detail> file:line" to stderr, then exit(code).  The checked
// expression is an mcs_* status (include/mcstar_hip.h).
#pragma once

#include <cstdio>
#include <cstdlib>

#include "../mcstar_hip.h"

namespace McsErr {
inline void check(int code, const char* file, int line) {
    if (code != MCS_OK) {
        fprintf(stderr, "MCSERR: %s (%s) %s:%d\n", mcs_error_string(code), mcs_last_error(), file, line);
        fflush(stderr);
        exit(code);
    }
}
}  // namespace McsErr

#define MCSERR(ans) McsErr::check((ans), __FILE__, __LINE__)
// the reference's macro name, for status checks in existing simulator code
#define CUERR(ans) McsErr::check((ans), __FILE__, __LINE__)
