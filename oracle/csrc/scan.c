/*
 * CPU restatement of the reference's Numba determinant-scan kernel
 * apply_operator_threaded (operator_state_algebra.py:139-219) -- TEST AND
 * CPU-BASELINE INFRASTRUCTURE ONLY (oracle).
 *
 * For one normal-ordered operator string: every target determinant that has
 * all creator bits set and no screened-annihilator bit set is mapped to its
 * source determinant by flipping the creator bits then the annihilator bits,
 * accumulating the Jordan-Wigner phase popcount(det & parity_check[orb]);
 * the source index comes from a hash map (the reference uses a Numba typed
 * dict), and out[i] += sign * factor * state[source].
 *
 * The hash map is open addressing with linear probing (keys are determinant
 * integers >= 0; empty slot = -1).
 */
#include <stdint.h>
#include <stdlib.h>

static inline uint64_t mix(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

void oracle_map_build(const int64_t* idx2det, int64_t n, int64_t* keys, int64_t* vals, int64_t cap) {
  for (int64_t i = 0; i < cap; ++i) keys[i] = -1;
  for (int64_t i = 0; i < n; ++i) {
    uint64_t h = mix((uint64_t)idx2det[i]) & (uint64_t)(cap - 1);
    while (keys[h] != -1) h = (h + 1) & (uint64_t)(cap - 1);
    keys[h] = idx2det[i];
    vals[h] = i;
  }
}

static inline int64_t map_get(const int64_t* keys, const int64_t* vals, int64_t cap, int64_t key) {
  uint64_t h = mix((uint64_t)key) & (uint64_t)(cap - 1);
  while (1) {
    int64_t k = keys[h];
    if (k == key) return vals[h];
    if (k == -1) return -1;
    h = (h + 1) & (uint64_t)(cap - 1);
  }
}

/* Returns 0, or -1 when a source determinant is outside the space (the
 * reference raises KeyError there unless do_unsafe). */
int oracle_apply_string(const double* state, double* out, const int64_t* idx2det, int64_t ndet, const int64_t* keys,
                        const int64_t* vals, int64_t cap, const int64_t* a_string, int len, int64_t create_mask,
                        int64_t anni_mask, const int64_t* parity_check, int nbits, double factor, int do_unsafe,
                        int nthreads) {
  int bad = 0;
  const int nb1 = nbits - 1;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(| : bad) if (nthreads > 1)
  for (int64_t i = 0; i < ndet; ++i) {
    int64_t det = idx2det[i];
    if ((det & create_mask) != create_mask) continue;
    if ((det & anni_mask) != 0) continue;
    int phase = 0;
    for (int k = 0; k < len; ++k) {
      int orb = (int)a_string[k];
      det ^= (int64_t)1 << (nb1 - orb);
      phase += __builtin_popcountll((uint64_t)(det & parity_check[orb]));
    }
    int64_t src = map_get(keys, vals, cap, det);
    if (src < 0) {
      if (!do_unsafe) bad = 1;
      continue;
    }
    double sign = 1.0 - 2.0 * (double)(phase & 1);
    out[i] += sign * factor * state[src];
  }
  return bad ? -1 : 0;
}
