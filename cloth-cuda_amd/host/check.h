#pragma once
// Error convention of the host layer: a failed libclothhip call prints the
// call site and the library's message, then exits, as the reference's
// CUDA_CHECK does (src/utils/cuda_helper.h:9-17).
#include "clothhip.h"

#include <cstdio>
#include <cstdlib>

#define CLH_CHECK(call)                                                        \
    do {                                                                       \
        const int _st = (call);                                                \
        if (_st != CLH_OK) {                                                   \
            std::fprintf(stderr, "clothhip error at %s:%d (status %d) - %s\n", \
                         __FILE__, __LINE__, _st, clh_last_error());           \
            std::exit(EXIT_FAILURE);                                           \
        }                                                                      \
    } while (0)

namespace cloth {
// Informational printing of the build steps (the reference prints the same
// summaries unconditionally); CLOTH_QUIET=1 in the environment silences it.
bool verbose();
}
