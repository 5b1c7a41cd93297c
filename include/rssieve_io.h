/*
 * rssieve_io.h -- host-side formats of the sieve, part of librssieve.so:
 * sr2sieve/sr5sieve ABCD input, the factor file, the stateA/B checkpoint,
 * and the pre-sieve that writes ABCD files.  Byte-compatible with the
 * reference (mfl0p/RSSieve v1.3); each entry point cites what it replaces.
 * Same conventions as rssieve.h: 0 or a negative RS_E_* code, messages via
 * rs_last_error(), no exit().
 */
#ifndef RSSIEVE_IO_H
#define RSSIEVE_IO_H

#include <stdint.h>

#include "rssieve.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rs_abcd rs_abcd;

/* Parse an ABCD file (inputfile.cpp:251-465): blank and '#' lines skipped,
 * headers "ABCD k*b^$a+-1 [n0]" with an optional " // Sieved to p", delta
 * lines accumulate n; one base, c = +-1, at most 100 sequences.  The bitmap
 * origin nmin is the smallest n rounded down to even.  On a format error the
 * message is the reference's "file:line: (seq i) msg" text. */
int rs_abcd_read(const char *path, rs_abcd **out);
void rs_abcd_free(rs_abcd *a);
int rs_abcd_info(const rs_abcd *a, uint32_t *base, int32_t *kcount, uint32_t *nmin, uint32_t *nmax,
                 uint64_t *words_per_seq, uint64_t *nterms);
/* k[i] and c[i] of every sequence (arrays of kcount entries). */
int rs_abcd_sequences(const rs_abcd *a, uint32_t *k, int32_t *c);
/* The candidate bitmap, kcount * words_per_seq words (inputfile.cpp:15-21). */
uint32_t *rs_abcd_bitmap(rs_abcd *a);
/* Clear the bit of k*b^n+c (sign '+' or '-') if it is a sequence of the file
 * (mark_factor_used, inputfile.cpp:89-107).  With k31_compat the sequence is
 * matched on the reference's signed-int encoding.  Returns 1 if cleared. */
int rs_abcd_mark_used(rs_abcd *a, uint32_t k, char sign, uint32_t n, int32_t k31_compat);

/* Write an ABCD file for sequences sharing one bitmap: one header per
 * sequence with at least one remaining term, then delta lines. */
int rs_abcd_write(const char *path, uint32_t base, int32_t kcount, const uint32_t *k, const int32_t *c, uint32_t nmin,
                  uint32_t nmax, const uint32_t *bitmap, uint64_t words_per_seq, uint64_t sieved_to);

/* Pre-sieve: clear in 'bitmap' every term k[i]*b^n+c[i], n in [nmin, nmax],
 * that has a prime factor p < bound (p >= 2), by discrete logarithms modulo
 * each p (baby-step giant-step in the cyclic group of b mod p).  Exact;
 * threaded over p.  bound <= 2^32. */
int rs_presieve(uint32_t base, int32_t kcount, const uint32_t *k, const int32_t *c, uint32_t nmin, uint32_t nmax,
                uint32_t *bitmap, uint64_t words_per_seq, uint64_t bound, int nthreads);

/* The factor-file line of cl_sieve.cpp:624: "p | k*b^n+c\n" with c printed
 * as %+d.  Returns the number of characters (excluding NUL) or RS_E_ARG. */
int rs_format_factor(char *buf, int buflen, uint64_t p, uint32_t k, uint32_t b, uint32_t n, int32_t c);
/* Parse one factor-file line as cl_sieve.cpp:797-863 does; 0 on success,
 * RS_E_IO if the line is not "p | k*b^n+-1" with base b. */
int rs_parse_factor(const char *line, uint32_t base, uint64_t *p, uint32_t *k, char *sign, uint32_t *n);

/* Checkpoint state: the reference's workStatus (cl_sieve.h:21-26), written
 * raw to stateA.ckp / stateB.ckp alternately with an additive checksum
 * (cl_sieve.cpp:185-328). */
typedef struct {
  uint64_t pmin, pmax, p, checksum, primecount, factorcount, last_trickle, state_sum, dupcount;
  uint32_t nmin, nmax, base;
  int32_t kcount;
  int32_t klist[RS_MAX_SEQUENCES];
} rs_state;

/* Write the state to dir/stateA.ckp or dir/stateB.ckp (*write_a_next picks
 * which; it flips after a successful write). */
int rs_state_write(const char *dir, rs_state *st, int *write_a_next);
/* Load the newest valid state whose base, pmin, pmax, nmin, nmax and kcount
 * match *st.  Returns 1 if resumed (and sets *write_a_next), 0 if no valid
 * file, negative on error. */
int rs_state_read(const char *dir, rs_state *st, int *write_a_next);

#ifdef __cplusplus
}
#endif
#endif /* RSSIEVE_IO_H */
