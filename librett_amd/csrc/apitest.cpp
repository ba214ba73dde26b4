// Links the C++ (mangled) librett.h API exactly as a LibreTT caller does and
// checks results on the host. Cases: the plain 64x48x32 float64 transpose (
// perm 2 0 1) and the known-answer case of the reference's tests/example.cpp
// (dims 3 3 3 2 4, perm 1 3 0 4 2) at 4, 8 and 16 bytes.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "librett.h"

static bool run_case(std::vector<int> dim, std::vector<int> perm, size_t elem, bool measure) {
  const int rank = (int)dim.size();
  size_t vol = 1;
  for (int d : dim) vol *= (size_t)d;
  std::vector<unsigned char> h_in(vol * elem), h_out(vol * elem, 0);
  for (size_t i = 0; i < vol * elem; i++) h_in[i] = (unsigned char)((i * 2654435761u) >> 13);
  void *d_in = nullptr, *d_out = nullptr;
  if (hipMalloc(&d_in, vol * elem) != hipSuccess || hipMalloc(&d_out, vol * elem) != hipSuccess) return false;
  hipStream_t stream;
  if (hipStreamCreate(&stream) != hipSuccess) return false;
  if (hipMemcpy(d_in, h_in.data(), vol * elem, hipMemcpyHostToDevice) != hipSuccess) return false;
  librettHandle plan;
  librettResult r = measure ? librettPlanMeasure(&plan, rank, dim.data(), perm.data(), elem, stream, d_in, d_out)
                            : librettPlan(&plan, rank, dim.data(), perm.data(), elem, stream);
  if (r != LIBRETT_SUCCESS) {
    std::printf("plan failed %d\n", (int)r);
    return false;
  }
  if (librettExecute(plan, d_in, d_out) != LIBRETT_SUCCESS) return false;
  if (hipStreamSynchronize(stream) != hipSuccess) return false;
  if (librettDestroy(plan) != LIBRETT_SUCCESS) return false;
  if (librettDestroy(plan) != LIBRETT_INVALID_PLAN) return false;
  if (hipMemcpy(h_out.data(), d_out, vol * elem, hipMemcpyDeviceToHost) != hipSuccess) return false;
  // Host check written like tests/example.cpp:175-212.
  std::vector<size_t> abc(rank), cba(rank);
  abc[0] = cba[0] = 1;
  for (int i = 1; i < rank; i++) {
    abc[i] = abc[i - 1] * dim[i - 1];
    cba[i] = cba[i - 1] * dim[perm[i - 1]];
  }
  std::vector<size_t> ijk(rank);
  bool ok = true;
  for (size_t index = 0; index < vol && ok; index++) {
    size_t rem = index;
    for (int r2 = rank - 1; r2 >= 0; r2--) {
      ijk[r2] = rem / abc[r2];
      rem -= abc[r2] * ijk[r2];
    }
    size_t t = 0;
    for (int r2 = 0; r2 < rank; r2++) t += cba[r2] * ijk[perm[r2]];
    ok = std::memcmp(&h_out[t * elem], &h_in[index * elem], elem) == 0;
  }
  (void)hipFree(d_in);
  (void)hipFree(d_out);
  (void)hipStreamDestroy(stream);
  std::printf("case rank %d elem %zu measure %d: %s\n", rank, elem, (int)measure, ok ? "ok" : "MISMATCH");
  return ok;
}

int main() {
  librettInitialize();
  bool ok = true;
  ok &= run_case({64, 48, 32}, {2, 0, 1}, 8, false);
  ok &= run_case({64, 48, 32}, {2, 0, 1}, 8, true);
  for (size_t e : {4, 8, 16}) ok &= run_case({3, 3, 3, 2, 4}, {1, 3, 0, 4, 2}, e, false);
  // Error codes (reference src/librett.cpp:85-108, :237, :377-379).
  int dim[2] = {4, 4}, bad[2] = {0, 0}, perm[2] = {1, 0};
  librettHandle h;
  hipStream_t s0 = 0;
  ok &= librettPlan(&h, 2, dim, bad, 8, s0) == LIBRETT_INVALID_PARAMETER;
  ok &= librettPlan(&h, 2, dim, perm, 2, s0) == LIBRETT_INVALID_PARAMETER;
  ok &= librettExecute(123456u, nullptr, nullptr) == LIBRETT_INVALID_PLAN;
  librettFinalize();
  std::printf(ok ? "apitest OK\n" : "apitest FAILED\n");
  return ok ? 0 : 1;
}
