/*
 * librett.h -- public API of the gfx950 (MI355X) tensor transpose library.
 *
 * Drop-in for LibreTT's header (reference src/librett.h:25-123): the same
 * include guard, the same handle and result types with the same values, and
 * the same six C++ entry points with identical parameter types, so code that
 * was compiled against LibreTT links against this library unchanged.
 *
 * Tensors are column-major (dim[0] varies fastest). The output has dims
 * dim[permutation[0]], dim[permutation[1]], ...; element (i_0, ..., i_{r-1})
 * of the input lands at output index (i_{perm[0]}, i_{perm[1]}, ...).
 * Elements are opaque 4-, 8- or 16-byte words and are copied bit for bit.
 *
 * This build is single-vendor: it always targets HIP on gfx950, so
 * LIBRETT_USES_HIP is defined here for code that tests it. Like the
 * reference (src/librett.h:31-33) the header includes <hip/hip_runtime.h>,
 * so callers that rely on that include keep compiling.
 */
#ifndef LIBRETT_H
#define LIBRETT_H

#include <stddef.h>
#include <hip/hip_runtime.h>

#ifndef LIBRETT_USES_HIP
#define LIBRETT_USES_HIP 1
#endif

/* Stream type of the API (reference src/librett.h:31-33). */
using librett_gpuStream_t = hipStream_t;

/* Handle of a stored plan (reference src/librett.h:40). */
typedef unsigned int librettHandle;

/* Result codes; values 0..5 in this order (reference src/librett.h:43-50). */
typedef enum librettResult_t {
  LIBRETT_SUCCESS,           /* success */
  LIBRETT_INVALID_PLAN,      /* unknown or destroyed handle */
  LIBRETT_INVALID_PARAMETER, /* invalid argument */
  LIBRETT_INVALID_DEVICE,    /* reserved: plan used on another device */
  LIBRETT_INTERNAL_ERROR,    /* planning or HIP runtime failure */
  LIBRETT_UNDEFINED_ERROR,   /* anything else */
} librettResult;

/* Library set-up and tear-down (reference src/librett.h:57,62). Both are
 * optional: every entry point initialises what it needs on first use. */
void librettInitialize();
void librettFinalize();

/*
 * Creates a plan chosen by the gfx950 cost model (reference
 * src/librett.h:78). rank >= 1, dim[i] >= 1, permutation is a permutation of
 * 0..rank-1, sizeofType is 4, 8 or 16, and prod(dim) < 2^31. The stream is
 * captured by value; every librettExecute of the plan runs on it.
 */
librettResult librettPlan(librettHandle* handle, int rank, int* dim, int* permutation, size_t sizeofType,
                          librett_gpuStream_t& stream);

/*
 * Creates a plan by timing every candidate on the caller's buffers (reference
 * src/librett.h:96-97). idata and odata are device buffers of prod(dim)
 * elements; odata is overwritten. idata == odata is an invalid parameter.
 */
librettResult librettPlanMeasure(librettHandle* handle, int rank, int* dim, int* permutation, size_t sizeofType,
                                 librett_gpuStream_t& stream, void* idata, void* odata);

/* Destroys a plan (reference src/librett.h:108). */
librettResult librettDestroy(librettHandle handle);

/*
 * Enqueues the out-of-place transpose on the plan's stream (reference
 * src/librett.h:121). Asynchronous: the caller synchronises the stream.
 */
librettResult librettExecute(librettHandle handle, void* idata, void* odata);

#endif /* LIBRETT_H */
