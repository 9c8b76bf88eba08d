// ORACLE harness (test infrastructure only): exposes rocRAND's own
// Philox4x32-10 rounds (rocrand/rocrand_philox4x32_10.h, a header shipped
// with ROCm) through a C ABI, as an independent implementation to check the
// oracle's and the kernel's Philox against.  Host code only.
#include <rocrand/rocrand_philox4x32_10.h>

namespace {
struct Rounds : rocrand_device::philox4x32_10_engine {
  uint4 apply(uint4 c, uint2 k) { return this->ten_rounds(c, k); }
};
}  // namespace

extern "C" void rocrand_philox4x32_10_rounds(const unsigned* ctr, const unsigned* key, unsigned* out) {
  Rounds r;
  uint4 c;
  c.x = ctr[0]; c.y = ctr[1]; c.z = ctr[2]; c.w = ctr[3];
  uint2 k;
  k.x = key[0]; k.y = key[1];
  uint4 o = r.apply(c, k);
  out[0] = o.x; out[1] = o.y; out[2] = o.z; out[3] = o.w;
}
