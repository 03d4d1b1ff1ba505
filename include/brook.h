/* brook_amd C ABI: the skeletonization path of giorgioangel/brook's C API on one AMD Instinct
 * MI355X (gfx950). Calls block; errors return a brook_status and brook_last_error() holds the
 * text (thread-local, valid until the next call on that thread). Inputs are borrowed for the call;
 * results are owned by their handles and views stay valid until the handle is destroyed.
 *
 * Each declaration names the reference interface it replaces (cpp/include/brook/brook.h:line).
 * Entry points marked EXTENSION have no counterpart in the reference's C API. */
#ifndef BROOK_AMD_BROOK_H
#define BROOK_AMD_BROOK_H

#include <stddef.h>
#include <stdint.h>

#include "dlpack.h"

#ifdef __cplusplus
extern "C" {
#endif

#define BROOK_ABI_VERSION 1u

typedef struct brook_context brook_context;
typedef struct brook_array brook_array;
typedef struct brook_components brook_components;
typedef struct brook_skeletons brook_skeletons;

/* brook.h:26-29. BROOK_CUDA_ERROR keeps the reference's name of value 3 (a device error: HIP
 * here), so consumer source that names it compiles unchanged. */
typedef enum brook_status {
  BROOK_SUCCESS = 0, BROOK_INVALID_ARGUMENT = 1, BROOK_OUT_OF_MEMORY = 2,
  BROOK_DEVICE_ERROR = 3, BROOK_CUDA_ERROR = 3, BROOK_INTERNAL_ERROR = 4
} brook_status;

/* brook.h:31-35 */
typedef enum brook_dtype {
  BROOK_U8 = 0, BROOK_U16 = 1, BROOK_U32 = 2, BROOK_U64 = 3,
  BROOK_I8 = 4, BROOK_I16 = 5, BROOK_I32 = 6, BROOK_I64 = 7,
  BROOK_F32 = 8, BROOK_F64 = 9
} brook_dtype;

/* brook.h:37 */
typedef enum brook_memory { BROOK_HOST = 0, BROOK_DEVICE = 1 } brook_memory;

/* brook.h:42-50: data points at logical (0,0,0); byte strides (any on the host; device input is
 * gathered into Fortran order on the GPU when it is not Fortran-contiguous). Zero-initialize and
 * set struct_size/abi_version. Device input must be on the context's device and ready. */
typedef struct brook_volume {
  size_t struct_size;
  uint32_t abi_version;
  const void *data;
  int64_t shape[3];
  int64_t strides[3];
  brook_dtype dtype;
  brook_memory memory;
} brook_volume;

/* brook.h:52-58 */
typedef struct brook_teasar_parameters {
  double scale, constant, pdrf_scale, pdrf_exponent;
  double soma_detection_threshold, soma_acceptance_threshold;
  double soma_invalidation_scale, soma_invalidation_const;
  int64_t max_paths;
  uint32_t has_max_paths;
} brook_teasar_parameters;

/* brook.h:60-67 */
enum brook_skeletonize_flags {
  BROOK_FIX_BRANCHING = 1u, BROOK_FIX_BORDERS = 2u, BROOK_FILTER_OBJECTS = 4u,
  BROOK_DEVICE_OUTPUT = 8u, BROOK_FILL_HOLES = 16u, BROOK_FIX_AVOCADOS = 32u
};

/* brook.h:68-84 */
typedef struct brook_skeletonize_options {
  size_t struct_size;
  uint32_t abi_version;
  brook_teasar_parameters teasar;
  float anisotropy[3];
  double dust_threshold;
  uint32_t flags;
  const int64_t *object_ids;
  size_t object_ids_count;
  const int64_t *targets_before; /* (x,y,z) int64 triples, caller order */
  size_t targets_before_count;
  const int64_t *targets_after;
  size_t targets_after_count;
  const brook_volume *voxel_graph;
} brook_skeletonize_options;

/* brook.h:88-94: host packed result; offsets have skeleton_count+1 entries; edges local. */
typedef struct brook_skeletons_view {
  size_t skeleton_count, vertex_count, edge_count;
  const int64_t *labels, *vertex_offsets, *edge_offsets;
  const float *vertices; /* vertex_count x 3, physical coordinates */
  const uint32_t *edges; /* edge_count x 2 */
  const float *radii;
  float anisotropy[3];
} brook_skeletons_view;

/* brook.h:147-150 */
uint32_t brook_abi_version(void);
const char *brook_version(void);
const char *brook_last_error(void);

/* brook.h:154-156: a negative or out-of-range device returns BROOK_INVALID_ARGUMENT. */
brook_status brook_context_create(int device, brook_context **out);
void brook_context_destroy(brook_context *context);
brook_status brook_context_synchronize(brook_context *context);

/* brook.h:159-162: device arrays owned by their handle. copy_to_host copies the array's bytes into
 * a host buffer of at least that many bytes. */
brook_status brook_array_upload(brook_context *context, const brook_volume *input, brook_array **out);
void brook_array_destroy(brook_array *array);
brook_status brook_array_view(const brook_array *array, brook_volume *out);
brook_status brook_array_copy_to_host(const brook_array *array, void *out, size_t bytes);
/* brook.h:163-167: zero-copy DLPack exports (device type kDLROCM, the array's device; shape and
 * element strides of the array's view: 3-D volumes, or the 2-D / 1-D rows of a skeleton buffer).
 * An export keeps the storage alive independently of the array and context handles; the consumer
 * calls its deleter exactly once. The versioned form reports version 1.min(max_minor,
 * DLPACK_MINOR_VERSION) and no flags. */
brook_status brook_array_to_dlpack(const brook_array *array, DLManagedTensor **out);
brook_status brook_array_to_dlpack_versioned(const brook_array *array, uint32_t max_minor,
                                             DLManagedTensorVersioned **out);

/* brook.h:169-182: 26-connected components, ids 1..N by each component's first voxel in
 * Fortran order, the smallest unsigned dtype that holds N; mapping has N+1 entries
 * (mapping[id] = input value), roots N (the first voxel of each id). */
brook_status brook_connected_components(brook_context *context, const brook_volume *labels, brook_components **out);
void brook_components_destroy(brook_components *components);
brook_status brook_components_view(const brook_components *components, brook_volume *labels,
                                   const int64_t **mapping, const int64_t **roots, size_t *count);
brook_status brook_components_copy_to_host(const brook_components *components, void *out, size_t bytes);
brook_status brook_components_labels(const brook_components *components, brook_array **out);

/* EXTENSION with no counterpart in the reference: brook_connected_components for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc); the call makes no device allocation. labels: device memory on the context's
 * device, Fortran-contiguous. ids_out: device memory of at least 4 bytes per voxel; receives the
 * ids as uint32 in Fortran order. mapping (capacity + 1 entries) and roots (capacity entries):
 * host arrays, as brook_components_view gives them. *count receives the component count; a count
 * above capacity returns BROOK_INVALID_ARGUMENT with only *count written. The working memory is a
 * static device array of 128 MiB per device, held from the first context on that device until the
 * process ends; a volume needing more (4 bytes per voxel plus 16 per table entry, rounded) is
 * rejected with BROOK_INVALID_ARGUMENT. Every call on one device shares that array, which keeps
 * the last call's contents: two calls on one GPU would race on it, so the library runs them one at
 * a time (a per-device lock; each call returns after its kernels finished). */
brook_status brook_connected_components_noalloc(brook_context *context, const brook_volume *labels, void *ids_out,
                                                size_t ids_bytes, int64_t *mapping, int64_t *roots, size_t capacity,
                                                size_t *count);

/* brook.h:183-184: multi-label anisotropic EDT (float32, Fortran order, background 0). */
brook_status brook_edt(brook_context *context, const brook_volume *labels, const float anisotropy[3], int black_border,
                       brook_array **out);

/* brook.h:192-196: components of a voxel graph (cc3d convention: bit b of a voxel's word permits
 * one of its 26 steps; uint32 words, or uint8 for one XY plane) whatever the labels, numbered in
 * order of each component's lowest voxel, background included; ids only on foreground voxels;
 * mapping[id] = the label where the id's last run starts; roots[id - 1] = its lowest foreground
 * voxel. brook_graph_edt: the graph-aware EDT (a blocked step is an interface half a voxel away)
 * of unsigned labels. */
brook_status brook_graph_components(brook_context *context, const brook_volume *labels, const brook_volume *voxel_graph,
                                    brook_components **out);
brook_status brook_graph_edt(brook_context *context, const brook_volume *labels, const brook_volume *voxel_graph,
                             const float anisotropy[3], int black_border, brook_array **out);
/* EXTENSION: the ids Kimimaro gives voxel-graph components (keys[id] = rank of the id's component
 * among all of them, 1-based; keys[0] = 0), count + 1 entries; NULL and 0 for plain components. */
brook_status brook_components_keys(const brook_components *components, const int64_t **keys, size_t *count);

/* brook.h:198-201 */
void brook_skeletonize_options_init(brook_skeletonize_options *options);
brook_status brook_skeletonize(brook_context *context, const brook_volume *labels,
                               const brook_skeletonize_options *options, brook_skeletons **out);

/* brook.h:230,252 */
brook_status brook_skeletons_get_view(const brook_skeletons *result, brook_skeletons_view *out);
void brook_skeletons_destroy(brook_skeletons *result);

/* brook.h:96-102, 232-233: results requested with BROOK_DEVICE_OUTPUT also hold their vertices
 * (vertex_count x 3 f32), edges (edge_count x 2 u32) and radii (f32) in device memory, row major;
 * labels and offsets stay host metadata, and get_view reads the host snapshot. The device views
 * describe rows: vertices shape (V, 3, 1), byte strides (12, 4, 12V); edges (E, 2, 1), (8, 4, 8E);
 * radii (V, 1, 1), (4, 4V, 4V). get_array shares the handle's storage (valid after the handle is
 * destroyed, until the array is) and views it the same way. */
typedef struct brook_skeletons_device_view {
  size_t skeleton_count, vertex_count, edge_count;
  const int64_t *labels, *vertex_offsets, *edge_offsets;
  brook_volume vertices, edges, radii;
  float anisotropy[3];
} brook_skeletons_device_view;
typedef enum brook_skeleton_buffer { BROOK_VERTICES = 0, BROOK_EDGES = 1, BROOK_RADII = 2 } brook_skeleton_buffer;
brook_status brook_skeletons_get_device_view(const brook_skeletons *result, brook_skeletons_device_view *out);
brook_status brook_skeletons_get_array(const brook_skeletons *result, brook_skeleton_buffer buffer, brook_array **out);

/* brook.h:234-237: one result from `count` fragments, fragment i offset by origins[3i..3i+2]
 * (count x 3 float32, NULL for zero offsets): one vertex per (label, position), the first
 * occurrence's radius, edges remapped without self-loops or duplicates, vertices on no edge
 * dropped, labels ascending (packed_merge.cu:88-166). The result is a device-output handle
 * (brook_skeletons_get_device_view), with its host snapshot for get_view. */
brook_status brook_merge_skeletons(brook_context *context, const brook_skeletons *const *fragments,
                                   const float *origins, size_t count, brook_skeletons **out);
/* EXTENSION: brook_merge_skeletons over borrowed views (packed results from any source): labels and
 * offsets in host memory; vertices, edges and radii in host memory or on the context's device. */
brook_status brook_merge_views(brook_context *context, const brook_skeletons_view *fragments, const float *origins,
                               size_t count, brook_skeletons **out);

/* brook.h:234-239: the canonical postprocess of a result on the device with the reference's tie
 * rules (post.cu:649-651): remove_dust(dust_threshold), remove_loops, join_components,
 * remove_ticks(tick_threshold), then vertices on no edge dropped. Labels keep their order. The
 * result is a device-output handle. */
brook_status brook_postprocess_canonical(brook_context *context, const brook_skeletons *input,
                                         double dust_threshold, double tick_threshold, brook_skeletons **out);
/* EXTENSION: the postprocess over a borrowed view (as brook_merge_views), whole or one step, as
 * the reference's Python binding offers them (Context.remove_dust, remove_loops, join_components,
 * remove_ticks, postprocess; post.cu:426-651): BROOK_POST_ALL (a = dust, b = ticks),
 * BROOK_POST_DUST (a), BROOK_POST_LOOPS, BROOK_POST_JOIN, BROOK_POST_TICKS (a), BROOK_POST_PRUNE.
 * An invalid packed graph is BROOK_INVALID_ARGUMENT. */
typedef enum {
  BROOK_POST_ALL = 0,
  BROOK_POST_DUST = 1,
  BROOK_POST_LOOPS = 2,
  BROOK_POST_JOIN = 3,
  BROOK_POST_TICKS = 4,
  BROOK_POST_PRUNE = 5
} brook_post_step;
brook_status brook_postprocess_views(brook_context *context, const brook_skeletons_view *input, int step, double a,
                                     double b, brook_skeletons **out);
/* EXTENSION (the reference's Context.skeleton_components, post.cu:421-425): per vertex of the
 * view, the least vertex index of its connected component; capacity >= vertex_count. */
brook_status brook_skeleton_components(brook_context *context, const brook_skeletons_view *input, int64_t *out,
                                       size_t capacity);

/* EXTENSION (the reference batches only in Python/C++, batch.hpp:20-22): skeletonize `count`
 * volumes in one call; out receives `count` result handles, each equal to a separate
 * brook_skeletonize of that volume. */
brook_status brook_skeletonize_batch(brook_context *context, const brook_volume *const *volumes, size_t count,
                                     const brook_skeletonize_options *options, brook_skeletons **out);
/* EXTENSION: brook_skeletonize_batch with one voxel graph per sample (graphs[i] NULL: none; graphs
 * NULL: none at all); options->voxel_graph must be NULL. */
brook_status brook_skeletonize_batch_graphs(brook_context *context, const brook_volume *const *volumes,
                                            const brook_volume *const *graphs, size_t count,
                                            const brook_skeletonize_options *options, brook_skeletons **out);

/* EXTENSION: the preamble's fields for stage-level checks: DBF, DAF and PDRF (before the roots
 * become rails) as n-voxel float32 host arrays (any may be NULL), and per-component tables of
 * table_capacity entries (component id index; NULL to skip): traced (0/1), root, target (flat
 * indices), dbf_max, max_daf. *count receives the component count. */
brook_status brook_prepare_fields(brook_context *context, const brook_volume *labels,
                                  const brook_skeletonize_options *options, float *dbf, float *daf, float *pdrf,
                                  size_t table_capacity, uint8_t *traced, int64_t *root, int64_t *target,
                                  float *dbf_max, float *max_daf, int64_t *count);

/* EXTENSION: the tracer's target orders (lockstep_order.cu:47-151) of a component volume, for
 * stage-level checks: for every nonzero id in ascending order, the id's voxels (flat Fortran
 * indices) by descending DAF (+-inf and +-0 as 0), ties by descending index, concatenated into
 * order (capacity entries; *count receives the foreground voxel count). ids: unsigned, at most 32
 * bits, host or device; daf: n float32 host values in Fortran order. */
brook_status brook_target_order(brook_context *context, const brook_volume *ids, const float *daf, int64_t *order,
                                size_t capacity, size_t *count);

/* EXTENSION: a timing/counter of the context's last call (e.g. "trace_ms", "paths"; with kernel
 * timing on, "k_<kernel>_ms" and "k_<kernel>_n" from events on the context's stream). */
brook_status brook_context_stat(const brook_context *context, const char *name, double *value);

/* EXTENSION (the reference's Python binding fences a producer stream the same way,
 * bindings.cpp:105-120): the context's later work waits for the work queued so far on `stream`
 * (a hipStream_t; 0 the null stream), with an event and no host wait. Device input produced on
 * another stream is then read only after it is ready. */
brook_status brook_context_wait_stream(brook_context *context, void *stream);

/* EXTENSION: time the engine's main kernels with HIP events on the context's stream (adds one
 * host wait per timed kernel). */
brook_status brook_context_set_kernel_timing(brook_context *context, int enable);

#ifdef __cplusplus
}
#endif
#endif
