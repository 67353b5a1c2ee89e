// Known-answer check of the Philox4x32-10 block function used for the device
// momenta and the Metropolis random numbers (su2_launch.hpp) against the
// published test vectors of Salmon, Moraes, Dror and Shaw, "Parallel random
// numbers: as easy as 1, 2, 3" (SC11; the Random123 distribution's kat_vectors):
//   counter 0, key 0;  counter and key all ones;  counter and key from pi.
// Prints the four output words per vector.
#include <cstdio>

#include "../../su2_amd/csrc/su2_launch.hpp"

int main() {
  const uint32_t in[3][6] = {
      {0u, 0u, 0u, 0u, 0u, 0u},
      {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
      {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u}};
  for (int v = 0; v < 3; ++v) {
    uint32_t c[4] = {in[v][0], in[v][1], in[v][2], in[v][3]};
    philox4x32_10(c, in[v][4], in[v][5]);
    printf("%08x %08x %08x %08x\n", c[0], c[1], c[2], c[3]);
  }
  return 0;
}
