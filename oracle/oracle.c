/*
 * CPU oracle for the LibreTT tensor permutation.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the library links, imports or calls
 * this file. Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py use it, and only as the checker (or, for the CPU baseline, as the
 * thing timed on the host), never as the product path.
 *
 * Semantics restated from the reference:
 *   - tests/example.cpp:175-212 (host nested-loop check): input is
 *     column-major with dims dim[0..rank-1] (dim[0] fastest); for input
 *     multi-index i, the output linear index is
 *         tindex = sum_r cba[r] * i[perm[r]],  cba[r] = prod_{s<r} dim[perm[s]]
 *     and out[tindex] = in[index].
 *   - src/TensorTester.cpp:201-220 (calcTensorConv): for output linear
 *     position o, the expected input position is
 *         sum_j ((o / c_j) % d_j) * ct_j
 *     with c_{perm[j]} = prod_{s<j} dim[perm[s]], d_j = dim[j],
 *     ct_j = prod_{s<j} dim[s].
 * Elements are opaque byte strings of 4, 8 or 16 bytes (src/librett.cpp:87).
 *
 * Functions:
 *   oracle_transpose          output-order odometer (follows the closed form)
 *   oracle_transpose_inorder  input-order loop written like tests/example.cpp
 *   oracle_transpose_decode   per-element decode like src/TensorTester.cpp
 *   oracle_transpose_omp      oracle_transpose split over OpenMP threads
 *   oracle_fill_uniform       seeded values in [-1,1) (splitmix64 counter hash)
 * Parity pins for each live in tests/test_oracle.py (numpy.transpose, the
 * example.cpp known answer, inverse-permutation round trips).
 */
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_MAX_RANK 64

static int check_args(int rank, const int* dim, const int* perm, size_t elem,
                      long long* vol_out) {
  if (rank < 1 || rank > ORACLE_MAX_RANK) return -1;
  if (elem != 4 && elem != 8 && elem != 16) return -1;
  int seen[ORACLE_MAX_RANK];
  memset(seen, 0, sizeof(seen));
  long long vol = 1;
  for (int r = 0; r < rank; r++) {
    if (dim[r] < 1) return -1;
    if (perm[r] < 0 || perm[r] >= rank || seen[perm[r]]++) return -1;
    vol *= dim[r];
  }
  *vol_out = vol;
  return 0;
}

/* Strides of the input (column-major) and of the input dim that each output
 * dim walks: istr_o[r] = input stride of dim perm[r]. */
static void strides(int rank, const int* dim, const int* perm, long long* istr,
                    long long* istr_o, int* odim) {
  long long s = 1;
  for (int r = 0; r < rank; r++) { istr[r] = s; s *= dim[r]; }
  for (int r = 0; r < rank; r++) { istr_o[r] = istr[perm[r]]; odim[r] = dim[perm[r]]; }
}

static void copy_elem(unsigned char* dst, const unsigned char* src, size_t elem) {
  memcpy(dst, src, elem);
}

/* Output-order odometer over [o0, o1): the output position o has multi-index
 * j (output dims), and reads input position sum_r j[r] * istr[perm[r]]. */
static void transpose_range(int rank, const long long* istr_o, const int* odim,
                            size_t elem, const unsigned char* in,
                            unsigned char* out, long long o0, long long o1) {
  if (o0 >= o1) return;
  int j[ORACLE_MAX_RANK];
  long long rem = o0, ipos = 0;
  for (int r = 0; r < rank; r++) {
    j[r] = (int)(rem % odim[r]);
    rem /= odim[r];
    ipos += (long long)j[r] * istr_o[r];
  }
  for (long long o = o0; o < o1; o++) {
    copy_elem(out + (size_t)o * elem, in + (size_t)ipos * elem, elem);
    /* advance the odometer */
    int r = 0;
    while (r < rank) {
      j[r]++;
      ipos += istr_o[r];
      if (j[r] < odim[r]) break;
      ipos -= (long long)odim[r] * istr_o[r];
      j[r] = 0;
      r++;
    }
  }
}

int oracle_transpose(int rank, const int* dim, const int* perm, size_t elem,
                     const void* in, void* out) {
  long long vol;
  if (check_args(rank, dim, perm, elem, &vol)) return -1;
  long long istr[ORACLE_MAX_RANK], istr_o[ORACLE_MAX_RANK];
  int odim[ORACLE_MAX_RANK];
  strides(rank, dim, perm, istr, istr_o, odim);
  transpose_range(rank, istr_o, odim, elem, (const unsigned char*)in,
                  (unsigned char*)out, 0, vol);
  return 0;
}

int oracle_transpose_omp(int rank, const int* dim, const int* perm, size_t elem,
                         const void* in, void* out, int nthreads) {
  long long vol;
  if (check_args(rank, dim, perm, elem, &vol)) return -1;
  long long istr[ORACLE_MAX_RANK], istr_o[ORACLE_MAX_RANK];
  int odim[ORACLE_MAX_RANK];
  strides(rank, dim, perm, istr, istr_o, odim);
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  /* Chunks of whole output cache lines per task keep first-touch and write
   * streams local to a thread. */
  long long nchunk = (long long)nthreads * 16;
  if (nchunk > vol) nchunk = vol;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (long long c = 0; c < nchunk; c++) {
    long long o0 = vol * c / nchunk, o1 = vol * (c + 1) / nchunk;
    transpose_range(rank, istr_o, odim, elem, (const unsigned char*)in,
                    (unsigned char*)out, o0, o1);
  }
#else
  (void)nthreads;
  transpose_range(rank, istr_o, odim, elem, (const unsigned char*)in,
                  (unsigned char*)out, 0, vol);
#endif
  return 0;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Written the way tests/example.cpp:175-212 computes it: walk the INPUT,
 * split the input linear index with abc[] and scatter with cba[]. */
int oracle_transpose_inorder(int rank, const int* dim, const int* perm,
                             size_t elem, const void* in, void* out) {
  long long vol;
  if (check_args(rank, dim, perm, elem, &vol)) return -1;
  long long abc[ORACLE_MAX_RANK], cba[ORACLE_MAX_RANK];
  long long ijk[ORACLE_MAX_RANK];
  abc[0] = 1;
  cba[0] = 1;
  for (int i = 1; i < rank; i++) {
    abc[i] = abc[i - 1] * dim[i - 1];
    cba[i] = cba[i - 1] * dim[perm[i - 1]];
  }
  const unsigned char* src = (const unsigned char*)in;
  unsigned char* dst = (unsigned char*)out;
  for (long long index = 0; index < vol; index++) {
    long long remainder = index;
    for (int r = rank - 1; r >= 0; r--) {
      ijk[r] = remainder / abc[r];
      remainder -= abc[r] * ijk[r];
    }
    long long tindex = 0;
    for (int r = 0; r < rank; r++) tindex += cba[r] * ijk[perm[r]];
    copy_elem(dst + (size_t)tindex * elem, src + (size_t)index * elem, elem);
  }
  return 0;
}

/* Written the way src/TensorTester.cpp:201-220 checks a result: for every
 * OUTPUT position, decode the expected input position with (c, d, ct). */
int oracle_transpose_decode(int rank, const int* dim, const int* perm,
                            size_t elem, const void* in, void* out) {
  long long vol;
  if (check_args(rank, dim, perm, elem, &vol)) return -1;
  long long c[ORACLE_MAX_RANK], d[ORACLE_MAX_RANK], ct[ORACLE_MAX_RANK];
  c[perm[0]] = 1;
  ct[0] = 1;
  d[0] = dim[0];
  for (int i = 1; i < rank; i++) {
    c[perm[i]] = c[perm[i - 1]] * dim[perm[i - 1]];
    d[i] = dim[i];
    ct[i] = ct[i - 1] * dim[i - 1];
  }
  const unsigned char* src = (const unsigned char*)in;
  unsigned char* dst = (unsigned char*)out;
  for (long long o = 0; o < vol; o++) {
    long long ref = 0;
    for (int j = 0; j < rank; j++) ref += ((o / c[j]) % d[j]) * ct[j];
    copy_elem(dst + (size_t)o * elem, src + (size_t)ref * elem, elem);
  }
  return 0;
}

/* splitmix64 output function (Steele, Lea and Flood, OOPSLA 2014), used as a
 * counter-based hash: draw k of stream `seed` is mix(seed*G + (k+1)*G) with
 * G = 0x9E3779B97F4A7C15. */
static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static uint64_t draw(uint64_t seed, uint64_t k) {
  const uint64_t G = 0x9E3779B97F4A7C15ULL;
  return mix64(seed * G + (k + 1) * G);
}

/* Values uniform in [-1,1): f32 from the top 24 bits, f64 from the top 53
 * bits, c128 as two consecutive f64 draws (elem = 4, 8, 16 bytes). */
int oracle_fill_uniform(void* buf, long long nelem, int elem, uint64_t seed) {
  if (elem == 4) {
    float* f = (float*)buf;
    for (long long i = 0; i < nelem; i++) {
      uint64_t u = draw(seed, (uint64_t)i);
      f[i] = (float)((double)(u >> 40) * (1.0 / 16777216.0) * 2.0 - 1.0);
    }
  } else if (elem == 8 || elem == 16) {
    double* f = (double*)buf;
    long long n = nelem * (elem / 8);
    for (long long i = 0; i < n; i++) {
      uint64_t u = draw(seed, (uint64_t)i);
      f[i] = (double)(u >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    }
  } else {
    return -1;
  }
  return 0;
}
