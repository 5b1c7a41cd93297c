// C-linkage entry points over the reference's own verify_factor.c, compiled
// from its source in place (see Makefile target 'ref').  Test infrastructure:
// only tests/ load the resulting oracle/_ref/libref_verify.so.
#include <cstdint>

uint64_t invert(uint64_t p);
uint64_t m_mul(uint64_t a, uint64_t b, uint64_t p, uint64_t q);
uint64_t add(uint64_t a, uint64_t b, uint64_t p);
bool isPrime(uint64_t p, uint64_t q, uint64_t one, uint64_t pmo, uint64_t two, uint64_t r2);
int verify_factor(uint64_t p, uint64_t k, uint32_t n, int32_t c, uint32_t base);

extern "C" {
uint64_t ref_invert(uint64_t p) { return invert(p); }
uint64_t ref_m_mul(uint64_t a, uint64_t b, uint64_t p, uint64_t q) { return m_mul(a, b, p, q); }
uint64_t ref_add(uint64_t a, uint64_t b, uint64_t p) { return add(a, b, p); }
int ref_verify_factor(uint64_t p, uint64_t k, uint32_t n, int32_t c, uint32_t base) {
  return verify_factor(p, k, n, c, base);
}
int ref_is_prime(uint64_t p) {
  // the same Montgomery constants verify_factor() derives (verify_factor.c:134-141)
  uint64_t q = invert(p);
  uint64_t one = (0 - p) % p;
  uint64_t pmo = p - one;
  uint64_t two = add(one, one, p);
  uint64_t t = add(two, two, p);
  for (int i = 0; i < 5; ++i) t = m_mul(t, t, p, q);
  return isPrime(p, q, one, pmo, two, t) ? 1 : 0;
}
}
