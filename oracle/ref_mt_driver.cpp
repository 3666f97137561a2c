// Test infrastructure: prints the reference's own MtRng64::getReal2() stream (compiled from
// src/vendor/mersennetwister.h where it lies in the reference checkout) for a seed, one value per
// line with 17 significant digits.  Usage: mt_ref SEED COUNT
#include <cstdio>
#include <cstdlib>

#include "mersennetwister.h"

int main(int argc, char** argv)
{
  if (argc < 3)
    return 2;
  MtRng64 rng;
  rng.init(strtoull(argv[1], nullptr, 10));
  const int n = atoi(argv[2]);
  for (int i = 0; i < n; i++)
    printf("%.17g\n", rng.getReal2());
  return 0;
}
