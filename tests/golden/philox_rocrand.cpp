// Prints Philox4x32-10 blocks computed by rocRAND's open-source engine (host-callable ten_rounds of
// rocrand_device::philox4x32_10_engine, exposed through a subclass) for fixed counters and keys.
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
struct Engine : rocrand_device::philox4x32_10_engine {
  uint4 block(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
  Engine eng;
  unsigned s = 12345u;
  auto next = [&]() { s = s * 1664525u + 1013904223u; return s; };
  printf("[\n");
  for (int i = 0; i < 64; ++i) {
    uint4 c = {next(), next(), next(), next()};
    if (i == 0) c = {0, 0, 0, 0};
    if (i == 1) c = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
    uint2 k = {next(), next()};
    if (i == 0) k = {0, 0};
    if (i == 1) k = {0xffffffffu, 0xffffffffu};
    uint4 r = eng.block(c, k);
    printf("  [[%u, %u, %u, %u], [%u, %u], [%u, %u, %u, %u]]%s\n", c.x, c.y, c.z, c.w, k.x, k.y, r.x, r.y, r.z, r.w,
           i == 63 ? "" : ",");
  }
  printf("]\n");
  return 0;
}
