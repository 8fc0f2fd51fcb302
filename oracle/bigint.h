/* Oracle-only arbitrary-precision integers (test infrastructure, never shipped).
 *
 * Sign-magnitude integers with 32-bit limbs and a fixed capacity.  They back the
 * exact rational geometry of the CPU oracle, standing in for the reference's
 * Boost.Multiprecision cpp_int (morsehgp3d/include/morsehgp3d/exact/integer.hpp:13).
 * Every operation checks the capacity and aborts loudly instead of wrapping.
 */
#ifndef MHGP_ORACLE_BIGINT_H
#define MHGP_ORACLE_BIGINT_H

#include <stdint.h>

#define BI_LIMBS 240 /* 7680 bits */

typedef struct {
  int32_t n;  /* used limbs; 0 means zero */
  int32_t s;  /* -1, 0, +1 */
  uint32_t d[BI_LIMBS];
} bi;

void bi_zero(bi *a);
void bi_set_i64(bi *a, int64_t v);
void bi_set_u64(bi *a, uint64_t v);
void bi_copy(bi *r, const bi *a);
void bi_add(bi *r, const bi *a, const bi *b);
void bi_sub(bi *r, const bi *a, const bi *b);
void bi_mul(bi *r, const bi *a, const bi *b);
void bi_mul_i64(bi *r, const bi *a, int64_t m);
void bi_shl(bi *r, const bi *a, int k);
void bi_neg(bi *r, const bi *a);
int bi_cmp(const bi *a, const bi *b);
int bi_cmpabs(const bi *a, const bi *b);
static inline int bi_sign(const bi *a) { return a->s; }
void bi_gcd(bi *r, const bi *a, const bi *b);
/* floor division of non-negative a by positive b; quotient and remainder */
void bi_divmod(bi *q, bi *rem, const bi *a, const bi *b);
int bi_to_dec(const bi *a, char *buf, int cap);
double bi_to_double(const bi *a);
int bi_bitlen(const bi *a);

#endif
