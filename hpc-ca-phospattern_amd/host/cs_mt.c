/*
 * Mersenne Twister MT19937 (Matsumoto & Nishimura, ACM TOMACS 8(1), 1998)
 * with the 2002 32-bit seeding, written from the published algorithm.
 *
 * The reference seeds its single global generator with mt_seed32new(-S)
 * (calcium_sparks.c:113-116, mtwist.c:407-439) and draws every gating
 * uniform with mt_ldrand (mtwist.h:635-677): two consecutive tempered
 * 32-bit outputs a, b combined as ((a << 32) | b) * 2^-64.  mtwist keeps
 * its state vector reversed, but the output stream equals the standard
 * init_genrand / genrand_int32 stream, which is what this file produces.
 */
#include <stdint.h>
#include "cs_host.h"

#define MT_N 624
#define MT_M 397
#define MT_MATRIX_A 0x9908b0dfu
#define MT_UPPER 0x80000000u
#define MT_LOWER 0x7fffffffu

void cs_mt_seed(cs_mt* g, uint32_t seed)
{
  g->mt[0] = seed;
  for (uint32_t i = 1; i < MT_N; i++) {
    uint32_t prev = g->mt[i - 1];
    g->mt[i] = 1812433253u * (prev ^ (prev >> 30)) + i;
  }
  g->idx = MT_N;
}

static void mt_twist(cs_mt* g)
{
  for (uint32_t k = 0; k < MT_N; k++) {
    uint32_t y = (g->mt[k] & MT_UPPER) | (g->mt[(k + 1) % MT_N] & MT_LOWER);
    uint32_t v = g->mt[(k + MT_M) % MT_N] ^ (y >> 1);
    if (y & 1u) v ^= MT_MATRIX_A;
    g->mt[k] = v;
  }
  g->idx = 0;
}

uint32_t cs_mt_u32(cs_mt* g)
{
  if (g->idx >= MT_N) mt_twist(g);
  uint32_t y = g->mt[g->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

double cs_mt_ldrand(cs_mt* g)
{
  uint64_t a = cs_mt_u32(g);
  uint64_t b = cs_mt_u32(g);
  /* 2^-64 is exact in binary; the product rounds the 64-bit integer to 53
   * bits first, as the reference's uint64 -> double conversion does. */
  return (double)((a << 32) | b) * 0x1p-64;
}
