/* A C consumer of include/brook.h, built against libbrook_amd.so by __graft_entry__.build() and
 * run by tests/test_consumers.py on a GPU. It checks the ABI conventions a C caller relies on, with
 * the reference's C test as the model (cpp/tests/c_smoke.c): known answers for components and the
 * EDT on an 8-voxel line, ownership of results and DLPack exports past their handles and context,
 * error statuses. Exit status 0 = pass; otherwise the number of the failed check. */
#include <stdio.h>
#include <string.h>

#include "brook.h"

#define REQUIRE(cond, code)                                        \
  do {                                                             \
    if (!(cond)) {                                                 \
      fprintf(stderr, "check %d failed: %s\n", code, brook_last_error()); \
      return code;                                                 \
    }                                                              \
  } while (0)
#define OK(call, code) REQUIRE((call) == BROOK_SUCCESS, code)

typedef char device_error_alias[(BROOK_CUDA_ERROR == BROOK_DEVICE_ERROR) ? 1 : -1];

static brook_volume line_of(const void *data, brook_dtype dtype, int bytes) {
  brook_volume v;
  memset(&v, 0, sizeof v);
  v.struct_size = sizeof v;
  v.abi_version = BROOK_ABI_VERSION;
  v.data = data;
  v.dtype = dtype;
  v.memory = BROOK_HOST;
  v.shape[0] = 8;
  v.shape[1] = v.shape[2] = 1;
  v.strides[0] = bytes;
  v.strides[1] = v.strides[2] = 8 * bytes;
  return v;
}

int main(void) {
  const uint8_t labels[8] = {1, 1, 0, 0, 0, 0, 2, 2};
  brook_volume input = line_of(labels, BROOK_U8, 1), view;
  brook_context *ctx = NULL;
  REQUIRE(brook_context_create(0, NULL) == BROOK_INVALID_ARGUMENT, 1);
  REQUIRE(brook_context_create(-1, &ctx) == BROOK_INVALID_ARGUMENT && ctx == NULL, 2);
  OK(brook_context_create(0, &ctx), 3);

  /* components: ids by first voxel, mapping to the input values, roots = first voxels */
  brook_components *cc = NULL;
  OK(brook_connected_components(ctx, &input, &cc), 4);
  const int64_t *mapping = NULL, *roots = NULL;
  size_t count = 0;
  OK(brook_components_view(cc, &view, &mapping, &roots, &count), 5);
  REQUIRE(count == 2 && mapping[1] == 1 && mapping[2] == 2 && roots[0] == 0 && roots[1] == 6, 6);
  REQUIRE(view.memory == BROOK_DEVICE && view.dtype == BROOK_U8 && view.shape[0] == 8 && view.strides[1] == 8, 7);
  uint8_t ids[16];
  REQUIRE(brook_components_copy_to_host(cc, ids, 7) == BROOK_INVALID_ARGUMENT, 8);
  OK(brook_components_copy_to_host(cc, ids, sizeof ids), 9); /* a larger buffer is fine */
  REQUIRE(memcmp(ids, labels, 8) == 0, 10);
  REQUIRE(brook_components_copy_to_host(NULL, ids, 8) == BROOK_INVALID_ARGUMENT, 11);
  brook_components_destroy(cc);

  /* EDT of the line (no black border): 2 1 0 0 0 0 1 2 */
  const float aniso[3] = {1, 1, 1};
  brook_array *distance = NULL;
  OK(brook_edt(ctx, &input, aniso, 0, &distance), 12);
  brook_context_destroy(ctx); /* results outlive their context */
  float d[8];
  OK(brook_array_copy_to_host(distance, d, sizeof d), 13);
  const float expected[8] = {2, 1, 0, 0, 0, 0, 1, 2};
  REQUIRE(memcmp(d, expected, sizeof d) == 0, 14);

  /* DLPack exports keep the storage after the array and context are gone */
  DLManagedTensor *exported = NULL;
  DLManagedTensorVersioned *versioned = NULL;
  OK(brook_array_to_dlpack(distance, &exported), 15);
  OK(brook_array_to_dlpack_versioned(distance, 0, &versioned), 16);
  REQUIRE(exported->dl_tensor.device.device_type == kDLROCM && exported->dl_tensor.device.device_id == 0, 17);
  REQUIRE(exported->dl_tensor.ndim == 3 && exported->dl_tensor.shape[0] == 8 && exported->dl_tensor.strides[0] == 1 &&
              exported->dl_tensor.strides[1] == 8 && exported->dl_tensor.dtype.code == kDLFloat &&
              exported->dl_tensor.dtype.bits == 32,
          18);
  REQUIRE(versioned->version.major == 1 && versioned->version.minor == 0 && versioned->flags == 0, 19);
  REQUIRE(versioned->dl_tensor.data == exported->dl_tensor.data, 20);
  brook_array_destroy(distance);
  OK(brook_context_create(0, &ctx), 21);
  brook_volume borrowed = input;
  borrowed.data = (const char *)exported->dl_tensor.data + exported->dl_tensor.byte_offset;
  borrowed.dtype = BROOK_F32;
  borrowed.memory = BROOK_DEVICE;
  for (int i = 0; i < 3; ++i) borrowed.strides[i] = exported->dl_tensor.strides[i] * 4;
  OK(brook_array_upload(ctx, &borrowed, &distance), 22);
  exported->deleter(exported);
  versioned->deleter(versioned);
  float copied[8];
  OK(brook_array_copy_to_host(distance, copied, sizeof copied), 23);
  REQUIRE(memcmp(copied, expected, sizeof copied) == 0, 24);
  brook_array_destroy(distance);

  /* a C-ordered host volume is uploaded in Fortran order */
  uint32_t c_order[3][4][5], fortran[60];
  for (int x = 0; x < 3; ++x)
    for (int y = 0; y < 4; ++y)
      for (int z = 0; z < 5; ++z) c_order[x][y][z] = (uint32_t)(100 * x + 10 * y + z);
  brook_volume cv = line_of(c_order, BROOK_U32, 4);
  cv.shape[0] = 3; cv.shape[1] = 4; cv.shape[2] = 5;
  cv.strides[0] = 80; cv.strides[1] = 20; cv.strides[2] = 4;
  brook_array *up = NULL;
  OK(brook_array_upload(ctx, &cv, &up), 25);
  OK(brook_array_copy_to_host(up, fortran, sizeof fortran), 26);
  for (int z = 0; z < 5; ++z)
    for (int y = 0; y < 4; ++y)
      for (int x = 0; x < 3; ++x) REQUIRE(fortran[x + 3 * (y + 4 * z)] == c_order[x][y][z], 27);
  OK(brook_array_view(up, &view), 28);
  REQUIRE(view.shape[0] == 3 && view.strides[0] == 4 && view.strides[1] == 12 && view.strides[2] == 48, 29);
  brook_array_destroy(up);

  /* invalid arguments come back as statuses with a message */
  REQUIRE(brook_edt(ctx, NULL, aniso, 0, &distance) == BROOK_INVALID_ARGUMENT && brook_last_error()[0], 30);
  brook_volume bad = input;
  bad.abi_version = 99;
  REQUIRE(brook_connected_components(ctx, &bad, &cc) == BROOK_INVALID_ARGUMENT, 31);
  brook_context_destroy(ctx);
  printf("c consumer ok\n");
  return 0;
}
