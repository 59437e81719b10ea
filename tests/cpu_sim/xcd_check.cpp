// Checks lfx::xcd_block on the host: for every grid size up to a bound, the
// renumbering is a bijection of [0, G), and the workgroups the dispatcher
// sends to one XCD (g % 8 == x) get consecutive numbers in dispatch order.
// Prints "ok" or the first failing grid size.
#include <cstdio>
#include <vector>

#include "../../lifted_amd/csrc/lfx_xcd.hpp"

int main() {
  for (unsigned G = 1; G <= 5000; ++G) {
    std::vector<int> seen(G, 0);
    for (unsigned g = 0; g < G; ++g) {
      const unsigned b = lfx::xcd_block(g, G);
      if (b >= G || seen[b]++) {
        std::printf("not a bijection at G=%u g=%u\n", G, g);
        return 1;
      }
      if (g >= 8 && b != lfx::xcd_block(g - 8, G) + 1) {
        std::printf("XCD range not contiguous at G=%u g=%u\n", G, g);
        return 1;
      }
    }
  }
  std::printf("ok\n");
  return 0;
}
