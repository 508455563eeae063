// The autotuner's cache as nq_solve reads it (the tuner itself is n_queens_amd/tune.py).
//
// A text file, $NQ_TUNE_CACHE or else $HOME/.cache/n_queens_amd/tune.txt, one entry per line
// (shown wrapped here):
//
//   <device arch> N=<N> preplace=<depth> split=<rows> block=<b> chunk=<c> push=<p>
//       waves_per_cu=<w> stack_extra=<e> [any further key=value fields, e.g. ms=<time>]
//
// '#' starts a comment.  The arch is the device's gcnArchName (for example
// gfx950:sramecc+:xnack-).  The last entry for (arch, N) wins.  An entry describes the DFS launch
// that was fastest for whole boards of size N: the depth of the claimed task list, the rows the
// kernel walks below each task, and the launch shape; stack_extra is the number of LDS frames
// beyond the minimum.  Counts never depend on it.  No reference counterpart: the reference fixes
// its launch at compile time (src/macro.h:25-35, src/n_queens_cuda.cu:164-167).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "nq.h"
#include "nq_internal.h"

namespace nq {

std::string tune_cache_path() {
    const char *p = getenv("NQ_TUNE_CACHE");
    if (p) return p;  // an empty value switches the cache off
    const char *home = getenv("HOME");
    if (!home || !*home) return "";
    return std::string(home) + "/.cache/n_queens_amd/tune.txt";
}

// Parse one entry line; false for comments, blank or malformed lines.
static bool parse_entry(const char *line, std::string *arch, TunedEntry *e) {
    while (*line == ' ' || *line == '\t') ++line;
    if (!*line || *line == '#' || *line == '\n') return false;
    char a[128];
    int n = 0;
    if (sscanf(line, "%127s%n", a, &n) != 1) return false;
    *arch = a;
    *e = TunedEntry{};
    int got = 0;
    const char *p = line + n;
    char key[64];
    long long val = 0;
    int m = 0;
    while (sscanf(p, " %63[A-Za-z_]=%lld%n", key, &val, &m) == 2) {
        p += m;
        const int v = static_cast<int>(val);
        if (!strcmp(key, "N")) e->N = v, got |= 1;
        else if (!strcmp(key, "preplace")) e->preplace = v, got |= 2;
        else if (!strcmp(key, "split")) e->split = v, got |= 4;
        else if (!strcmp(key, "block")) e->block = v, got |= 8;
        else if (!strcmp(key, "chunk")) e->chunk = v, got |= 16;
        else if (!strcmp(key, "push")) e->push = v, got |= 32;
        else if (!strcmp(key, "waves_per_cu")) e->waves_per_cu = v, got |= 64;
        else if (!strcmp(key, "stack_extra")) e->stack_extra = v, got |= 128;
        // other keys (timings) are informative
        while (*p && *p != ' ' && *p != '\t' && *p != '\n') ++p;
    }
    if (got != 255) return false;
    return e->N >= 1 && e->N <= 31 && e->preplace >= 0 && e->preplace <= e->N && e->split >= 0 && e->split <= 4 &&
           e->block >= 0 && e->block <= 256 && e->block % 64 == 0 && e->chunk >= 0 && e->chunk <= 4096 &&
           e->push >= 0 && e->push <= 2 && e->waves_per_cu >= 0 && e->stack_extra >= 0 && e->stack_extra <= 8;
}

bool find_tuned(const char *arch, int N, TunedEntry *out) {
    const std::string path = tune_cache_path();
    if (path.empty() || !arch) return false;
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return false;
    char line[512];
    bool found = false;
    std::string a;
    TunedEntry e;
    while (fgets(line, sizeof line, f)) {
        if (parse_entry(line, &a, &e) && a == arch && e.N == N) {
            *out = e;
            found = true;
        }
    }
    fclose(f);
    return found;
}

}  // namespace nq
