/*
 * Exhaustive check of the order-0 rANS encoder's quotient (test helper,
 * compiled by tests/test_capi.py with gcc).
 *
 * The encoder divides the renormalised state x by the symbol frequency f as
 * q = trunc(RN(double(x) * rc)), rc = 1/f rounded to nearest and raised by
 * one ulp when rc * f < 1 (so rc >= 1/f).  The state before a step
 * satisfies x < f * 2^19 (byte-wise rANS, L = 2^23, 12-bit scale), and
 * q(x) is non-decreasing in x, so q(x) == floor(x / f) for every such x as
 * soon as q(k f - 1) == k - 1 and q(k f) == k for every k in 1..2^19: each
 * interval [k f, (k+1) f - 1] is then pinned at both ends.  x86-64 double
 * arithmetic rounds to nearest like the GPU's v_mul_f64.
 */
#include <math.h>
#include <stdint.h>

static double reciprocal_up(uint32_t f) {
  double rc = 1.0 / (double)f;
  if (fma(rc, (double)f, -1.0) < 0.0) rc = nextafter(rc, INFINITY);
  return rc;
}

static uint32_t quotient(uint32_t x, double rc) { return (uint32_t)((double)x * rc); }

/* Returns 0 if the identity holds for every f in [f_lo, f_hi], else the
 * first failing f; *bad_x receives the failing state. */
int64_t check_rans_reciprocal(uint32_t f_lo, uint32_t f_hi, uint64_t* bad_x,
                              uint64_t* checked) {
  uint64_t n = 0;
  for (uint32_t f = f_lo; f <= f_hi; ++f) {
    const double rc = reciprocal_up(f);
    if (quotient(0, rc) != 0) { *bad_x = 0; *checked = n; return f; }
    for (uint32_t k = 1; k <= (1u << 19); ++k) {
      const uint32_t x = k * f;  /* <= 2^31 */
      if (quotient(x - 1, rc) != k - 1) { *bad_x = x - 1; *checked = n; return f; }
      if (k < (1u << 19) && quotient(x, rc) != k) { *bad_x = x; *checked = n; return f; }
      n += 2;
    }
  }
  *checked = n;
  return 0;
}

/* The same with a deliberately wrong reciprocal (rounded to nearest only,
 * never raised): the check must catch it for some f, or it proves nothing. */
int64_t check_rans_reciprocal_nearest(uint32_t f_lo, uint32_t f_hi, uint64_t* bad_x) {
  for (uint32_t f = f_lo; f <= f_hi; ++f) {
    const double rc = 1.0 / (double)f;
    for (uint32_t k = 1; k <= (1u << 19); ++k) {
      const uint32_t x = k * f;
      if (quotient(x - 1, rc) != k - 1 || (k < (1u << 19) && quotient(x, rc) != k)) {
        *bad_x = x;
        return f;
      }
    }
  }
  return 0;
}
