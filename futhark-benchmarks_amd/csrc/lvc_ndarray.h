/* The multi-dimensional arrays of include/futhark_arrays.h (not exported). */
#ifndef LVC_NDARRAY_H
#define LVC_NDARRAY_H

#include <stddef.h>
#include <stdint.h>

/* Row-major device array of rank <= 3. */
typedef struct {
  void *mem;
  int64_t shape[3];
  size_t bytes;
  int raw; /* wraps caller memory: not freed */
} lvc_ndarray;

struct futhark_f32_2d {
  lvc_ndarray a;
};
struct futhark_i32_2d {
  lvc_ndarray a;
};
struct futhark_f32_3d {
  lvc_ndarray a;
};

#endif /* LVC_NDARRAY_H */
