// Host build of the library's decimal parser (gahhg_amd/csrc/gahhg_decimal.h)
// for the tests: one number per input line; prints "OK <bits hex> <chars
// consumed>" or "FAIL".
#include <stdio.h>
#include <string.h>

#include "gahhg_decimal.h"

int main() {
  static char line[4096];
  while (fgets(line, sizeof line, stdin)) {
    int64_t n = (int64_t)strcspn(line, "\n");
    int64_t i = 0;
    double v = 0.0;
    if (gahhg::dec::parse_decimal((const uint8_t*)line, i, n, v)) {
      uint64_t b;
      memcpy(&b, &v, 8);
      printf("OK %016llx %lld\n", (unsigned long long)b, (long long)i);
    } else {
      printf("FAIL\n");
    }
  }
  return 0;
}
