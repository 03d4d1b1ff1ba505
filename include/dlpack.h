/* The DLPack tensor exchange ABI (the open DLPack specification, version 1.x): the structures
 * brook_array_to_dlpack and brook_array_to_dlpack_versioned fill in. Written for this library from
 * the published specification; the layouts, names and values are the standard's, so a consumer
 * that includes another copy of the official header first gets the same definitions (the include
 * guard is the official header's). Only the core exchange types are declared here. */
#ifndef DLPACK_DLPACK_H_
#define DLPACK_DLPACK_H_

#include <stddef.h>
#include <stdint.h>

#define DLPACK_MAJOR_VERSION 1
#define DLPACK_MINOR_VERSION 1

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint32_t major;
  uint32_t minor;
} DLPackVersion;

typedef enum {
  kDLCPU = 1,
  kDLCUDA = 2,
  kDLCUDAHost = 3,
  kDLOpenCL = 4,
  kDLVulkan = 7,
  kDLMetal = 8,
  kDLVPI = 9,
  kDLROCM = 10,
  kDLROCMHost = 11,
  kDLExtDev = 12,
  kDLCUDAManaged = 13,
  kDLOneAPI = 14,
  kDLWebGPU = 15,
  kDLHexagon = 16,
  kDLMAIA = 17
} DLDeviceType;

typedef struct {
  DLDeviceType device_type;
  int32_t device_id;
} DLDevice;

typedef enum {
  kDLInt = 0U,
  kDLUInt = 1U,
  kDLFloat = 2U,
  kDLOpaqueHandle = 3U,
  kDLBfloat = 4U,
  kDLComplex = 5U,
  kDLBool = 6U
} DLDataTypeCode;

typedef struct {
  uint8_t code;
  uint8_t bits;
  uint16_t lanes;
} DLDataType;

/* strides and shape in elements; strides may be NULL for a compact row-major tensor. */
typedef struct {
  void *data;
  DLDevice device;
  int32_t ndim;
  DLDataType dtype;
  int64_t *shape;
  int64_t *strides;
  uint64_t byte_offset;
} DLTensor;

/* The consumer calls deleter(self) exactly once when it no longer needs the tensor. */
typedef struct DLManagedTensor {
  DLTensor dl_tensor;
  void *manager_ctx;
  void (*deleter)(struct DLManagedTensor *self);
} DLManagedTensor;

#define DLPACK_FLAG_BITMASK_READ_ONLY (1UL << 0UL)
#define DLPACK_FLAG_BITMASK_IS_COPIED (1UL << 1UL)

typedef struct DLManagedTensorVersioned {
  DLPackVersion version;
  void *manager_ctx;
  void (*deleter)(struct DLManagedTensorVersioned *self);
  uint64_t flags;
  DLTensor dl_tensor;
} DLManagedTensorVersioned;

#ifdef __cplusplus
}
#endif
#endif /* DLPACK_DLPACK_H_ */
