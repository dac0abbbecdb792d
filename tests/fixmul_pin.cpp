// Exposes the oracle's fix_mul_f32 (o * N exactly, truncated toward zero to the fixed point's
// 2^-64) and its volume-term products (acc_prod: x*y, acc_triple: x*y*z) to tests/test_fixmul.py,
// which checks them against exact rational arithmetic.
#include <cstdint>

#include "../oracle/onum.hpp"

extern "C" void fixmul_pin(int64_t nHi, uint64_t nLo, float o, int64_t* rHi, uint64_t* rLo) {
  const orc::i128 N = (orc::i128)(((unsigned __int128)(uint64_t)nHi << 64) | nLo);
  const orc::i128 r = orc::fix_mul_f32(N, o);
  *rHi = (int64_t)(r >> 64);
  *rLo = (uint64_t)r;
}

extern "C" void prod_pin(float x, float y, int neg, int64_t* rHi, uint64_t* rLo) {
  orc::i128 acc = 0;
  orc::acc_prod(acc, x, y, neg != 0);
  *rHi = (int64_t)(acc >> 64);
  *rLo = (uint64_t)acc;
}

extern "C" void triple_pin(float x, float y, float z, int neg, int64_t* rHi, uint64_t* rLo) {
  orc::i128 acc = 0;
  orc::acc_triple(acc, x, y, z, neg != 0);
  *rHi = (int64_t)(acc >> 64);
  *rLo = (uint64_t)acc;
}
