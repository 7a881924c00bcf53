// Leveled logging: note() prints when level >= the current log level.
// Same contract as src/note.c:7-20 (default level 1, printf semantics).
#include <cstdarg>
#include <cstdio>

#include "host_common.h"

namespace {
unsigned int g_log_level = 1;
}

extern "C" void set_log_level(unsigned int level) { g_log_level = level; }

extern "C" void note(unsigned int level, const char *format, ...) {
    if (level < g_log_level) return;
    va_list args;
    va_start(args, format);
    std::vprintf(format, args);
    va_end(args);
    std::fflush(stdout);
}
