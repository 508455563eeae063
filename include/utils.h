// utils.h — the reference's two logging helpers (src/utils.h:5-7), exported by libnq_hip.so with
// the same signatures and C++ linkage, so that the reference's src/main.cpp (which calls them at
// src/main.cpp:49-54) links against libnq_hip.so alone.
#pragma once

#include <sys/time.h>

// Milliseconds from start to end (src/utils.cpp:10-12).
float time_diff_ms(struct timeval &start, struct timeval &end);

// printf to stdout behind a "[YYYY-mm-dd HH:MM:SS.mmm] " local-time stamp (src/utils.cpp:14-30).
void print_with_time(const char *format, ...);
