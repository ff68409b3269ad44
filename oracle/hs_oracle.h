/*
 * hs_oracle.h -- CPU restatement of the reference's variable-time MSM path, used ONLY as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  It is test
 * infrastructure, never linked into or called by the product library (helioselene_amd/).
 *
 * Curve index `c`: 0 = Helios (coordinates in F_p, p = 2^255-19, group order q),
 *                  1 = Selene (coordinates in F_q, q = 2^255-gamma, group order p).
 * All byte strings are 32-byte little-endian; points are the reference's compressed encoding
 * (x LE, bit 255 = parity of canonical y, identity = 0^32: helios/src/x64/helios_tobytes.cpp:42-66).
 * On input the point decoders map 0^32 to the identity like the reference test loader
 * (src/tests.cpp:7479-7496) and otherwise follow helios_frombytes (helios_frombytes.cpp:59-109).
 */
#ifndef HS_ORACLE_H
#define HS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* field k: 0 = F_p, 1 = F_q. a,b,out: canonical 32-byte LE (inputs reduced mod the prime). */
void hso_fe_mul(int k, uint8_t out[32], const uint8_t a[32], const uint8_t b[32]);
void hso_fe_sq(int k, uint8_t out[32], const uint8_t a[32]);
void hso_fe_add(int k, uint8_t out[32], const uint8_t a[32], const uint8_t b[32]);
void hso_fe_sub(int k, uint8_t out[32], const uint8_t a[32], const uint8_t b[32]);
void hso_fe_inv(int k, uint8_t out[32], const uint8_t a[32]);
int hso_fe_sqrt(int k, uint8_t out[32], const uint8_t a[32]);
void hso_batch_invert(int k, uint8_t *out, const uint8_t *in, size_t n);

/* curve c */
int hso_point_add(int c, uint8_t out[32], const uint8_t a[32], const uint8_t b[32]);
int hso_point_dbl(int c, uint8_t out[32], const uint8_t a[32]);
int hso_point_neg(int c, uint8_t out[32], const uint8_t a[32]);
int hso_point_valid(int c, const uint8_t a[32]);
int hso_scalar_mul(int c, uint8_t out[32], const uint8_t s[32], const uint8_t p[32]);
void hso_generator(int c, uint8_t out[32]);
void hso_reduce_wide(int c, uint8_t out[32], const uint8_t wide[64]);
void hso_map_to_curve(int c, uint8_t *out, const uint8_t *u, size_t n);
void hso_map_to_curve2(int c, uint8_t out[32], const uint8_t u0[32], const uint8_t u1[32]);

/* MSM as the reference computes it (Straus n<=16, Pippenger otherwise, including its window-8
 * top-digit behaviour).  Returns -1 if a point fails to decode. */
int hso_msm(int c, uint8_t out[32], const uint8_t *scalars, const uint8_t *points, size_t n);
/* Exact sum of s_i * P_i (double-and-add per term), the reference tests' naive check. */
int hso_msm_naive(int c, uint8_t out[32], const uint8_t *scalars, const uint8_t *points, size_t n);
/* r*H + sum v_i G_i as one MSM of n+1 terms (helios/include/helios_pedersen.h:55-86). */
int hso_pedersen(int c, uint8_t out[32], const uint8_t blinding[32], const uint8_t H[32],
                 const uint8_t *values, const uint8_t *gens, size_t n);
/* Fixed-base MSM over 16-entry affine tables (helios/include/helios_msm_fixed.h:60-155). */
int hso_msm_fixed(int c, uint8_t out[32], const uint8_t *scalars, const uint8_t *points, size_t n);
/* The 16-entry affine table [1P..16P] of one base, as 16 x (x[32] || y[32]) canonical LE. */
int hso_fixed_table(int c, uint8_t out[16 * 64], const uint8_t p[32]);

/* Window width the reference picks for n (helios_msm_vartime.cpp:248-263), 0 when n <= 16. */
int hso_ref_window(size_t n);
/* 1 if the reference's width-w signed recode drops a final carry for this scalar. */
int hso_recode_drops_carry(const uint8_t s[32], int w);
/* Reference signed recode (encode_signed_wbit, helios_msm_vartime.cpp:127-162). Returns count. */
int hso_recode(int16_t *digits, const uint8_t s[32], int w);

/* xoshiro256** with splitmix64 seeding (src/tests_fuzz.cpp:103-157). */
void hso_xoshiro_fill(uint64_t seed, uint8_t *buf, size_t len);

/* Affine coordinates (x || y, 64 bytes canonical LE) of a compressed point; 0 ok, -1 invalid,
 * 1 identity (out zeroed). */
int hso_decompress(int c, uint8_t out[64], const uint8_t p[32]);

#ifdef __cplusplus
}
#endif
#endif
