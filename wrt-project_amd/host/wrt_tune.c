/*
 * wrt_tune.c -- the tuning cache shared by the per-job brick launch (band shape, direction
 * group) and the batch planner (direction group): text lines "<signature hex> <v0> <v1>
 * <seconds>", appended as tunings finish, the last line of a signature winning.  The file
 * is WRT_TUNE_CACHE, else $HOME/.cache/wrt_tune.txt (else /tmp/wrt_tune.txt); without a
 * writable file tuning is simply redone in the next process.  WRT_AUTOTUNE=0 turns tuning
 * off (the defaults are used).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "wrt_internal.h"

static void cache_path(char* buf, size_t len) {
  const char* e = getenv("WRT_TUNE_CACHE");
  if (e && *e) {
    snprintf(buf, len, "%s", e);
    return;
  }
  const char* home = getenv("HOME");
  if (home && *home) {
    snprintf(buf, len, "%s/.cache", home);
    mkdir(buf, 0755); /* fails harmlessly when it exists or cannot be made */
    snprintf(buf, len, "%s/.cache/wrt_tune.txt", home);
  } else {
    snprintf(buf, len, "/tmp/wrt_tune.txt");
  }
}

int wrt_tune_enabled(void) {
  const char* at = getenv("WRT_AUTOTUNE");
  return !(at && at[0] == '0');
}

unsigned long long wrt_tune_hash(const long long* v, int n) {
  unsigned long long h = 1469598103934665603ULL; /* FNV-1a over the words */
  for (int i = 0; i < n; ++i) {
    h ^= (unsigned long long)v[i];
    h *= 1099511628211ULL;
  }
  return h;
}

int wrt_tune_lookup(unsigned long long sig, int* v0, int* v1) {
  char path[1024], line[256];
  cache_path(path, sizeof(path));
  FILE* f = fopen(path, "r");
  if (!f) return 0;
  int found = 0;
  while (fgets(line, sizeof(line), f)) {
    unsigned long long s;
    int a, b;
    double sec;
    if (sscanf(line, "%llx %d %d %lf", &s, &a, &b, &sec) == 4 && s == sig) {
      *v0 = a;
      *v1 = b;
      found = 1;
    }
  }
  fclose(f);
  return found;
}

void wrt_tune_store(unsigned long long sig, int v0, int v1, double sec) {
  char path[1024];
  cache_path(path, sizeof(path));
  FILE* f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%016llx %d %d %.6f\n", sig, v0, v1, sec);
  fclose(f);
}
