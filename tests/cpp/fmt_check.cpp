// Host build of the device formatter (differential-equations_amd/csrc/de_fmt.h)
// for the CPU tests: reads doubles (raw little-endian bytes) on stdin, writes
// one `{:?}` text per line and, with -c, the number of texts whose digit
// search reached a loop cap.  Built with -fsanitize=address,undefined.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../differential-equations_amd/csrc/de_fmt.h"

int main(int argc, char** argv) {
  const bool count_only = argc > 1 && strcmp(argv[1], "-c") == 0;
  std::vector<double> xs;
  double x;
  while (fread(&x, sizeof x, 1, stdin) == 1) xs.push_back(x);
  long bad = 0;
  char buf[de::ryu::kMaxChars + 1];
  for (double v : xs) {
    const de::FmtF64 f = de::fmt_prepare(v);
    if (f.len < 1 || f.len > de::ryu::kMaxChars) return 2;
    bad += f.bad;
    de::fmt_emit(f, buf);
    buf[f.len] = 0;
    if (!count_only) printf("%s\n", buf);
  }
  if (count_only) printf("%ld\n", bad);
  return 0;
}
