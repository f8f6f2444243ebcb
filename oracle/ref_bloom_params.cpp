// Driver that prints hash_parameter(nc) of the reference's own
// auto_tune_bloom.h (compiled where it lies) for nc = 1..4096, one line per
// nc: "nc hash_len bit hash".  Used only to pin the oracle's restatement.
#include "auto_tune_bloom.h"

int main() {
  for (int nc = 1; nc <= 4096; ++nc) {
    int hash_len = -1, bit = -1, hash = -1;
    hash_parameter(nc, hash_len, bit, hash);
    printf("%d %d %d %d\n", nc, hash_len, bit, hash);
  }
  return 0;
}
