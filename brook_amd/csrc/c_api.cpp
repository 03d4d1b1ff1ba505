// The C ABI (include/brook.h): handles, error protection and input normalization around the
// engine (the conventions of the reference's c_api.cpp:17-41, 218-277).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/brook.h"
#include "engine.hpp"
#include "graph.hpp"
#include "layout.hpp"
#include "merge.hpp"
#include "post.hpp"

namespace {

// Device memory owned by a result handle, independent of the context that made it.
struct DeviceMemory {
  void *ptr = nullptr;
  size_t bytes = 0;
  int device = 0;
  DeviceMemory(int dev, size_t n) : bytes(n), device(dev) {
    if (n) BROOK_HIP(hipMalloc(&ptr, n));
  }
  ~DeviceMemory() {
    if (ptr) { (void)hipSetDevice(device); (void)hipFree(ptr); }
  }
};

thread_local std::string g_error;

template <class F> brook_status protect(F &&fn) noexcept {
  try {
    g_error.clear();
    fn();
    return BROOK_SUCCESS;
  } catch (const std::invalid_argument &e) {
    g_error = e.what();
    return BROOK_INVALID_ARGUMENT;
  } catch (const std::bad_alloc &e) {
    g_error = e.what();
    return BROOK_OUT_OF_MEMORY;
  } catch (const brook::DeviceError &e) {
    g_error = e.what();
    return e.code == hipErrorOutOfMemory ? BROOK_OUT_OF_MEMORY : BROOK_DEVICE_ERROR;
  } catch (const std::exception &e) {
    g_error = e.what();
    return BROOK_INTERNAL_ERROR;
  } catch (...) {
    g_error = "unknown exception";
    return BROOK_INTERNAL_ERROR;
  }
}

void require(bool ok, const char *what = "null or invalid brook argument") {
  if (!ok) throw std::invalid_argument(what);
}

int dtype_bytes(brook_dtype t) {
  switch (t) {
    case BROOK_U8: case BROOK_I8: return 1;
    case BROOK_U16: case BROOK_I16: return 2;
    case BROOK_U32: case BROOK_I32: case BROOK_F32: return 4;
    case BROOK_U64: case BROOK_I64: case BROOK_F64: return 8;
  }
  throw std::invalid_argument("invalid dtype");
}
bool integer_dtype(brook_dtype t) { return t >= BROOK_U8 && t <= BROOK_I64; }

void check_volume(const brook_volume *v) {
  require(v != nullptr);
  require(v->struct_size >= sizeof(brook_volume) && v->abi_version == BROOK_ABI_VERSION, "incompatible brook_volume");
  for (int a = 0; a < 3; ++a) require(v->shape[a] >= 0, "negative volume shape");
  int64_t n = v->shape[0] * v->shape[1] * v->shape[2];
  require(n == 0 || v->data != nullptr, "null volume data");
  require(n < (int64_t(1) << 31) - 4096, "volume exceeds the int32 voxel index space");
  require(v->memory == BROOK_HOST || v->memory == BROOK_DEVICE, "invalid memory kind");
}

// Whether the strides lay the volume out densely from `data` upwards in some axis order (every
// byte of the n * b span belongs to one voxel, positive strides).
bool dense(const brook_volume *v, int b) {
  int axes[3] = {0, 1, 2};
  std::sort(axes, axes + 3, [&](int p, int q) { return v->strides[p] < v->strides[q]; });
  int64_t pitch = b;
  for (int a : axes) {
    if (v->shape[a] <= 1) continue;
    if (v->strides[a] != pitch) return false;
    pitch *= v->shape[a];
  }
  return true;
}

// The volume as Fortran-ordered device memory: borrowed when it already is, else copied.
struct Input {
  brook::LabelVolume view;
  std::unique_ptr<brook::Buffer<uint8_t>> holder;
};

Input normalize(brook::Context &ctx, const brook_volume *v) {
  check_volume(v);
  Input in;
  const int b = dtype_bytes(v->dtype);
  const int64_t sx = v->shape[0], sy = v->shape[1], sz = v->shape[2], n = sx * sy * sz;
  in.view.bytes = b;
  in.view.is_signed = v->dtype >= BROOK_I8 && v->dtype <= BROOK_I64;
  in.view.dims = brook::Dims{int(sx), int(sy), int(sz)};
  const bool fortran = (sx <= 1 || v->strides[0] == b) && (sy <= 1 || v->strides[1] == b * sx) &&
                       (sz <= 1 || v->strides[2] == b * sx * sy);
  if (n == 0) return in;
  if (v->memory == BROOK_DEVICE && fortran) {
    in.view.data = v->data;
    return in;
  }
  in.holder = std::make_unique<brook::Buffer<uint8_t>>(ctx, size_t(n) * b);
  const long long strides[3] = {v->strides[0], v->strides[1], v->strides[2]};
  if (v->memory == BROOK_DEVICE) {
    brook::gather_fortran(ctx, v->data, strides, b, in.view.dims, in.holder->data());
  } else if (fortran) {
    BROOK_HIP(hipMemcpyAsync(in.holder->data(), v->data, size_t(n) * b, hipMemcpyHostToDevice, ctx.stream()));
  } else if (dense(v, b)) {
    // a dense host layout in another axis order (C order included): its bytes go up as they are
    // and are gathered into Fortran order on the GPU (layout.cu:33-61)
    brook::Buffer<uint8_t> staged(ctx, size_t(n) * b);
    BROOK_HIP(hipMemcpyAsync(staged.data(), v->data, size_t(n) * b, hipMemcpyHostToDevice, ctx.stream()));
    brook::gather_fortran(ctx, staged.data(), strides, b, in.view.dims, in.holder->data());
    ctx.synchronize();
  } else {
    std::vector<uint8_t> packed(size_t(n) * b);
    const char *base = static_cast<const char *>(v->data);
    size_t o = 0;
    for (int64_t z = 0; z < sz; ++z)
      for (int64_t y = 0; y < sy; ++y)
        for (int64_t x = 0; x < sx; ++x, o += b)
          std::memcpy(&packed[o], base + x * v->strides[0] + y * v->strides[1] + z * v->strides[2], b);
    BROOK_HIP(hipMemcpyAsync(in.holder->data(), packed.data(), packed.size(), hipMemcpyHostToDevice, ctx.stream()));
    ctx.synchronize();
  }
  in.view.data = in.holder->data();
  return in;
}

brook_dtype unsigned_dtype(int bytes) { return bytes == 1 ? BROOK_U8 : bytes == 2 ? BROOK_U16 : bytes == 4 ? BROOK_U32 : BROOK_U64; }

// A voxel graph as one uint32 word per voxel in Fortran device memory (graph_components.cu:178-182:
// uint32, or uint8 for one XY plane, whose values are widened as they are).
struct GraphInput {
  Input in;
  std::unique_ptr<brook::Buffer<unsigned>> wide;
  const unsigned *words = nullptr;
  bool compact_2d = false;
};

GraphInput normalize_graph(brook::Context &ctx, const brook_volume *g, const brook_volume *labels) {
  check_volume(g);
  bool same = true;
  for (int a = 0; a < 3; ++a) same = same && g->shape[a] == labels->shape[a];
  require(same && (g->dtype == BROOK_U32 || (g->dtype == BROOK_U8 && g->shape[2] == 1)),
          "voxel_graph must match the label shape and use uint32, or uint8 for an XY plane");
  GraphInput gi;
  gi.in = normalize(ctx, g);
  gi.compact_2d = g->dtype == BROOK_U8;
  const int64_t n = gi.in.view.size();
  if (gi.compact_2d && n) {
    gi.wide = std::make_unique<brook::Buffer<unsigned>>(ctx, size_t(n));
    brook::widen_words(ctx, static_cast<const uint8_t *>(gi.in.view.data), n, gi.wide->data());
    gi.words = gi.wide->data();
  } else {
    gi.words = static_cast<const unsigned *>(gi.in.view.data);
  }
  return gi;
}

}  // namespace

struct brook_context {
  brook::Context value;
  explicit brook_context(int device) : value(device) {}
};
// A device array: a volume (shape, Fortran order), or the rows of a skeleton buffer (rows x
// width, row major; width 1: a vector), as the reference's tensor views (packed.cpp:7-14).
struct brook_array {
  std::shared_ptr<DeviceMemory> memory;
  brook_dtype dtype = BROOK_U8;
  int64_t shape[3] = {0, 0, 0};
  int width = 0;  // > 0: rows of `width` elements (shape[0] rows)
};
struct brook_components {
  brook_array labels;
  std::vector<int64_t> mapping, roots;
  std::vector<int64_t> keys;  // voxel-graph components: Kimimaro's id of each id (empty otherwise)
};
struct brook_skeletons {
  brook::Packed packed;
  // BROOK_DEVICE_OUTPUT: device copies of vertices (V x 3 f32), edges (E x 2 u32) and radii (V f32)
  std::shared_ptr<DeviceMemory> vertices, edges, radii;
};

namespace {

// The logical dimensions of an array: element counts and byte strides (ndim entries; the view's
// remaining axes have extent 1 and the array's byte size as stride, as the reference's
// Array::view, runtime.cpp:81-95).
struct Dims3 {
  int ndim = 3;
  int64_t shape[3] = {0, 0, 0}, strides[3] = {0, 0, 0};
};
Dims3 logical(const brook_array &a) {
  const int64_t b = dtype_bytes(a.dtype);
  Dims3 d;
  if (a.width > 0) {
    d.ndim = a.width == 1 ? 1 : 2;
    d.shape[0] = a.shape[0];
    d.strides[0] = b * a.width;
    d.shape[1] = a.width;
    d.strides[1] = b;
  } else {
    int64_t pitch = b;
    for (int i = 0; i < 3; ++i) { d.shape[i] = a.shape[i]; d.strides[i] = pitch; pitch *= a.shape[i]; }
  }
  return d;
}
brook_volume view_of_array(const brook_array &a) {
  brook_volume out{};
  out.struct_size = sizeof(brook_volume);
  out.abi_version = BROOK_ABI_VERSION;
  out.data = a.memory->ptr;
  out.dtype = a.dtype;
  out.memory = BROOK_DEVICE;
  const Dims3 d = logical(a);
  for (int i = 0; i < 3; ++i) {
    out.shape[i] = i < d.ndim ? d.shape[i] : 1;
    out.strides[i] = i < d.ndim ? d.strides[i] : int64_t(a.memory->bytes);
  }
  return out;
}
// Rows of `width` elements in device memory a skeleton result owns, as a shared array.
brook_array rows_array(const std::shared_ptr<DeviceMemory> &m, brook_dtype dtype, int width, int64_t rows) {
  brook_array a;
  a.memory = m;
  a.dtype = dtype;
  a.width = width;
  a.shape[0] = rows;
  a.shape[1] = width;
  a.shape[2] = 1;
  return a;
}

// DLPack export of an array: the tensor owns a reference to the storage (brook.h:163-167).
template <class Tensor> struct Export {
  Tensor tensor{};
  std::shared_ptr<DeviceMemory> storage;
  int64_t shape[3] = {0, 0, 0}, strides[3] = {0, 0, 0};
};
template <class Tensor> void release_export(Tensor *t) noexcept { delete static_cast<Export<Tensor> *>(t->manager_ctx); }
template <class Tensor> Tensor *export_array(const brook_array &a) {
  auto e = std::make_unique<Export<Tensor>>();
  e->storage = a.memory;
  const Dims3 d = logical(a);
  const int64_t b = dtype_bytes(a.dtype);
  for (int i = 0; i < d.ndim; ++i) { e->shape[i] = d.shape[i]; e->strides[i] = d.strides[i] / b; }
  e->tensor.manager_ctx = e.get();
  e->tensor.deleter = release_export<Tensor>;
  DLTensor &t = e->tensor.dl_tensor;
  t.data = a.memory->ptr;
  t.device = DLDevice{kDLROCM, int32_t(a.memory->device)};
  t.ndim = d.ndim;
  const bool is_float = a.dtype == BROOK_F32 || a.dtype == BROOK_F64;
  const bool is_signed = a.dtype >= BROOK_I8 && a.dtype <= BROOK_I64;
  t.dtype = DLDataType{uint8_t(is_float ? kDLFloat : is_signed ? kDLInt : kDLUInt), uint8_t(b * 8), 1};
  t.shape = e->shape;
  t.strides = e->strides;
  t.byte_offset = 0;
  return &e.release()->tensor;
}

brook_array *owned_array(brook::Context &ctx, const void *src, size_t bytes, brook_dtype dtype, const int64_t shape[3]) {
  auto a = std::make_unique<brook_array>();
  a->memory = std::make_shared<DeviceMemory>(ctx.device(), bytes);
  if (bytes) BROOK_HIP(hipMemcpyAsync(a->memory->ptr, src, bytes, hipMemcpyDeviceToDevice, ctx.stream()));
  ctx.synchronize();
  a->dtype = dtype;
  std::copy(shape, shape + 3, a->shape);
  return a.release();
}

brook::SkeletonizeOptions convert(const brook_skeletonize_options *p, std::vector<uint64_t> *ids) {
  require(p != nullptr);
  require(p->struct_size >= sizeof(brook_skeletonize_options) && p->abi_version == BROOK_ABI_VERSION,
          "incompatible skeletonization options");
  require(!(p->flags & ~uint32_t(BROOK_FIX_BRANCHING | BROOK_FIX_BORDERS | BROOK_FILTER_OBJECTS | BROOK_DEVICE_OUTPUT |
                                 BROOK_FILL_HOLES | BROOK_FIX_AVOCADOS)),
          "unknown skeletonization flags");
  brook::SkeletonizeOptions o;
  auto &t = o.teasar;
  t.scale = p->teasar.scale;
  t.constant = p->teasar.constant;
  t.pdrf_scale = p->teasar.pdrf_scale;
  t.pdrf_exponent = p->teasar.pdrf_exponent;
  t.soma_detection = p->teasar.soma_detection_threshold;
  t.soma_acceptance = p->teasar.soma_acceptance_threshold;
  t.soma_scale = p->teasar.soma_invalidation_scale;
  t.soma_constant = p->teasar.soma_invalidation_const;
  t.max_paths = p->teasar.has_max_paths ? p->teasar.max_paths : -1;
  std::copy(p->anisotropy, p->anisotropy + 3, o.aniso);
  o.dust = p->dust_threshold;
  o.fix_branching = p->flags & BROOK_FIX_BRANCHING;
  o.fix_borders = p->flags & BROOK_FIX_BORDERS;
  o.fill_holes = p->flags & BROOK_FILL_HOLES;
  o.fix_avocados = p->flags & BROOK_FIX_AVOCADOS;
  o.voxel_graph = p->voxel_graph != nullptr;
  auto points = [](const int64_t *data, size_t count) {
    require(data || !count);
    std::vector<brook::Point> out;
    for (size_t i = 0; i < count; ++i) out.push_back({data[3 * i], data[3 * i + 1], data[3 * i + 2]});
    return out;
  };
  o.before = points(p->targets_before, p->targets_before_count);
  o.after = points(p->targets_after, p->targets_after_count);
  if (ids) {
    ids->clear();
    if (p->flags & BROOK_FILTER_OBJECTS) {
      require(p->object_ids || !p->object_ids_count);
      for (size_t i = 0; i < p->object_ids_count; ++i) ids->push_back(uint64_t(p->object_ids[i]));
    }
  }
  return o;
}

// A caller's volume as the engine's device label volume (owned copies, if any, live in `in` and
// `masked`), with the object filter applied.
brook::LabelVolume engine_input(brook::Context &ctx, const brook_volume *v, const brook_skeletonize_options *p,
                                const std::vector<uint64_t> &ids, Input &in, brook::Buffer<uint8_t> &masked) {
  require(integer_dtype(v ? v->dtype : BROOK_F32), "skeletonize requires integer labels");
  in = normalize(ctx, v);
  if (p->flags & BROOK_FILTER_OBJECTS) {
    masked = brook::Buffer<uint8_t>(ctx, size_t(in.view.size()) * in.view.bytes);
    brook::mask_labels(ctx, in.view.data, in.view.bytes, in.view.size(), ids, masked.data());
    in.view.data = masked.data();
  }
  return in.view;
}

brook::Packed run_one(brook::Context &ctx, const brook_volume *v, const brook_skeletonize_options *p,
                      brook::FieldsOut *fields = nullptr) {
  std::vector<uint64_t> ids;
  brook::SkeletonizeOptions o = convert(p, &ids);
  Input in;
  brook::Buffer<uint8_t> masked;
  brook::LabelVolume view = engine_input(ctx, v, p, ids, in, masked);
  GraphInput graph;
  if (p->voxel_graph) {
    graph = normalize_graph(ctx, p->voxel_graph, v);
    view.graph = graph.words;
    view.graph_compact_2d = graph.compact_2d;
  }
  return brook::skeletonize(ctx, view, o, fields);
}

}  // namespace

extern "C" {

uint32_t brook_abi_version(void) { return BROOK_ABI_VERSION; }
const char *brook_version(void) { return "0.1.0+mi355x"; }
const char *brook_last_error(void) { return g_error.c_str(); }

brook_status brook_context_create(int device, brook_context **out) {
  return protect([&] {
    require(out != nullptr);
    *out = nullptr;
    require(device >= 0, "negative device index");
    auto ctx = std::make_unique<brook_context>(device);
    brook::noalloc_scratch_init(device);  // the device is current: its no-allocation scratch address
    *out = ctx.release();
  });
}
void brook_context_destroy(brook_context *context) { delete context; }
brook_status brook_context_synchronize(brook_context *context) {
  return protect([&] {
    require(context != nullptr);
    brook::ContextCall call(context->value);
    context->value.synchronize();
  });
}

brook_status brook_array_upload(brook_context *ctx, const brook_volume *input, brook_array **out) {
  return protect([&] {
    require(ctx && input && out);
    *out = nullptr;
    brook::ContextCall call(ctx->value);
    Input in = normalize(ctx->value, input);
    *out = owned_array(ctx->value, in.view.data, size_t(in.view.size()) * in.view.bytes, input->dtype, input->shape);
  });
}
void brook_array_destroy(brook_array *array) { delete array; }
brook_status brook_array_view(const brook_array *a, brook_volume *out) {
  return protect([&] {
    require(a && out);
    *out = view_of_array(*a);
  });
}
brook_status brook_array_copy_to_host(const brook_array *a, void *out, size_t bytes) {
  return protect([&] {
    require(a != nullptr);
    const size_t size = a->memory->bytes;
    require(bytes >= size && (out || !size), "output buffer is too small");
    int previous = 0;
    BROOK_HIP(hipGetDevice(&previous));
    BROOK_HIP(hipSetDevice(a->memory->device));
    const hipError_t e = size ? hipMemcpy(out, a->memory->ptr, size, hipMemcpyDeviceToHost) : hipSuccess;
    (void)hipSetDevice(previous);
    BROOK_HIP(e);
  });
}
brook_status brook_array_to_dlpack(const brook_array *a, DLManagedTensor **out) {
  return protect([&] {
    require(a && out);
    *out = nullptr;
    *out = export_array<DLManagedTensor>(*a);
  });
}
brook_status brook_array_to_dlpack_versioned(const brook_array *a, uint32_t max_minor, DLManagedTensorVersioned **out) {
  return protect([&] {
    require(a && out);
    *out = nullptr;
    DLManagedTensorVersioned *t = export_array<DLManagedTensorVersioned>(*a);
    t->version = DLPackVersion{1, std::min<uint32_t>(max_minor, DLPACK_MINOR_VERSION)};
    t->flags = 0;
    *out = t;
  });
}

brook_status brook_connected_components(brook_context *ctx, const brook_volume *labels, brook_components **out) {
  return protect([&] {
    require(ctx && labels && out);
    *out = nullptr;
    require(integer_dtype(labels->dtype), "connected components requires integer labels");
    brook::ContextCall call(ctx->value);
    Input in = normalize(ctx->value, labels);
    brook::Components cc = brook::connected_components(ctx->value, in.view);
    auto r = std::make_unique<brook_components>();
    const size_t n = size_t(in.view.size());
    brook_dtype dt = unsigned_dtype(cc.cc_bytes);
    if (n) {
      auto *a = owned_array(ctx->value, cc.labels.data(), n * cc.cc_bytes, dt, labels->shape);
      r->labels = std::move(*a);
      delete a;
    } else {
      r->labels.memory = std::make_shared<DeviceMemory>(ctx->value.device(), 0);
      r->labels.dtype = BROOK_U32;
      std::copy(labels->shape, labels->shape + 3, r->labels.shape);
    }
    r->mapping = cc.mapping;
    r->roots = cc.roots;
    *out = r.release();
  });
}
void brook_components_destroy(brook_components *c) { delete c; }
brook_status brook_components_view(const brook_components *c, brook_volume *labels, const int64_t **mapping,
                                   const int64_t **roots, size_t *count) {
  return protect([&] {
    require(c != nullptr);
    if (labels) {
      brook_status s = brook_array_view(&c->labels, labels);
      if (s != BROOK_SUCCESS) throw std::runtime_error(g_error);
    }
    if (mapping) *mapping = c->mapping.data();
    if (roots) *roots = c->roots.data();
    if (count) *count = c->roots.size();
  });
}
brook_status brook_components_copy_to_host(const brook_components *c, void *out, size_t bytes) {
  if (!c) return protect([] { require(false); });
  return brook_array_copy_to_host(&c->labels, out, bytes);
}
brook_status brook_components_labels(const brook_components *c, brook_array **out) {
  return protect([&] {
    require(c && out);
    *out = new brook_array(c->labels);
  });
}

brook_status brook_connected_components_noalloc(brook_context *ctx, const brook_volume *labels, void *ids_out,
                                                size_t ids_bytes, int64_t *mapping, int64_t *roots, size_t capacity,
                                                size_t *count) {
  return protect([&] {
    require(ctx && labels && count && mapping && (roots || !capacity));
    *count = 0;
    check_volume(labels);
    require(integer_dtype(labels->dtype), "connected components requires integer labels");
    require(labels->memory == BROOK_DEVICE, "the no-allocation variant takes device labels");
    const int b = dtype_bytes(labels->dtype);
    const int64_t sx = labels->shape[0], sy = labels->shape[1], sz = labels->shape[2], n = sx * sy * sz;
    const bool fortran = (sx <= 1 || labels->strides[0] == b) && (sy <= 1 || labels->strides[1] == b * sx) &&
                         (sz <= 1 || labels->strides[2] == b * sx * sy);
    require(fortran, "the no-allocation variant takes Fortran-contiguous labels");
    require(ids_bytes >= size_t(n) * 4 && (ids_out || !n), "ids buffer smaller than 4 bytes per voxel");
    require(capacity < (size_t(1) << 40), "table capacity out of range");
    brook::ContextCall call(ctx->value);
    brook::LabelVolume view;
    view.data = labels->data;
    view.bytes = b;
    view.is_signed = labels->dtype >= BROOK_I8 && labels->dtype <= BROOK_I64;
    view.dims = brook::Dims{int(sx), int(sy), int(sz)};
    const int64_t got = brook::connected_components_noalloc(ctx->value, view, static_cast<uint32_t *>(ids_out), mapping,
                                                            roots, int64_t(capacity));
    *count = size_t(got);
    require(got <= int64_t(capacity), "component count exceeds the table capacity");
  });
}

brook_status brook_graph_components(brook_context *ctx, const brook_volume *labels, const brook_volume *voxel_graph,
                                   brook_components **out) {
  return protect([&] {
    require(ctx && labels && voxel_graph && out);
    *out = nullptr;
    require(integer_dtype(labels->dtype), "voxel graph labels must be integers");
    brook::ContextCall call(ctx->value);
    Input in = normalize(ctx->value, labels);
    GraphInput graph = normalize_graph(ctx->value, voxel_graph, labels);
    brook::GraphComponents g = brook::graph_components(ctx->value, in.view, graph.words, graph.compact_2d);
    auto r = std::make_unique<brook_components>();
    const size_t n = size_t(in.view.size());
    if (n) {
      auto *a = owned_array(ctx->value, g.cc.labels.data(), n * g.cc.cc_bytes, unsigned_dtype(g.cc.cc_bytes), labels->shape);
      r->labels = std::move(*a);
      delete a;
    } else {
      r->labels.memory = std::make_shared<DeviceMemory>(ctx->value.device(), 0);
      r->labels.dtype = BROOK_U8;
      std::copy(labels->shape, labels->shape + 3, r->labels.shape);
    }
    r->mapping = g.cc.mapping;
    r->roots = g.cc.roots;
    r->keys = g.keys;
    *out = r.release();
  });
}

brook_status brook_components_keys(const brook_components *c, const int64_t **keys, size_t *count) {
  return protect([&] {
    require(c && keys && count);
    *keys = c->keys.empty() ? nullptr : c->keys.data();
    *count = c->keys.size();
  });
}

brook_status brook_graph_edt(brook_context *ctx, const brook_volume *labels, const brook_volume *voxel_graph,
                             const float anisotropy[3], int black_border, brook_array **out) {
  return protect([&] {
    require(ctx && labels && voxel_graph && anisotropy && out);
    *out = nullptr;
    require(labels->dtype >= BROOK_U8 && labels->dtype <= BROOK_U64, "graph EDT expects unsigned component labels");
    brook::ContextCall call(ctx->value);
    Input in = normalize(ctx->value, labels);
    GraphInput graph = normalize_graph(ctx->value, voxel_graph, labels);
    const size_t n = size_t(in.view.size());
    auto a = std::make_unique<brook_array>();
    a->memory = std::make_shared<DeviceMemory>(ctx->value.device(), n * 4);
    a->dtype = BROOK_F32;
    std::copy(labels->shape, labels->shape + 3, a->shape);
    if (n) {
      brook::graph_edt(ctx->value, in.view.data, in.view.bytes, in.view.dims, graph.words, anisotropy, black_border != 0,
                       static_cast<float *>(a->memory->ptr));
      ctx->value.synchronize();
    }
    *out = a.release();
  });
}

brook_status brook_edt(brook_context *ctx, const brook_volume *labels, const float anisotropy[3], int black_border,
                       brook_array **out) {
  return protect([&] {
    require(ctx && labels && anisotropy && out);
    *out = nullptr;
    require(integer_dtype(labels->dtype), "EDT requires integer labels");
    brook::ContextCall call(ctx->value);
    Input in = normalize(ctx->value, labels);
    const size_t n = size_t(in.view.size());
    auto a = std::make_unique<brook_array>();
    a->memory = std::make_shared<DeviceMemory>(ctx->value.device(), n * 4);
    a->dtype = BROOK_F32;
    std::copy(labels->shape, labels->shape + 3, a->shape);
    if (n) {
      // Runs of equal value along an axis lie in one component, so the EDT of the labels equals
      // the EDT of their component ids; 64-bit labels go through the ids.
      if (in.view.bytes == 8) {
        brook::Components cc = brook::connected_components(ctx->value, in.view);
        brook::edt(ctx->value, cc.labels.data(), cc.cc_bytes, in.view.dims, anisotropy, black_border != 0, 3,
                   static_cast<float *>(a->memory->ptr));
        ctx->value.synchronize();
      } else {
        brook::edt(ctx->value, in.view.data, in.view.bytes, in.view.dims, anisotropy, black_border != 0, 3,
                   static_cast<float *>(a->memory->ptr));
      }
      ctx->value.synchronize();
    }
    *out = a.release();
  });
}

void brook_skeletonize_options_init(brook_skeletonize_options *p) {
  if (!p) return;
  *p = brook_skeletonize_options{};
  p->struct_size = sizeof(*p);
  p->abi_version = BROOK_ABI_VERSION;
  p->teasar.scale = 1.5;
  p->teasar.constant = 300;
  p->teasar.pdrf_scale = 100000;
  p->teasar.pdrf_exponent = 4;
  p->teasar.soma_detection_threshold = 750;
  p->teasar.soma_acceptance_threshold = 3500;
  p->teasar.soma_invalidation_scale = 2;
  p->teasar.soma_invalidation_const = 300;
  p->anisotropy[0] = p->anisotropy[1] = p->anisotropy[2] = 1;
  p->dust_threshold = 1000;
  p->flags = BROOK_FIX_BRANCHING | BROOK_FIX_BORDERS;
}

namespace {
// BROOK_DEVICE_OUTPUT (brook.h:66, c_api.cpp:314-353): the packed arrays also live in device memory
// the handle owns; the host arrays stay as the handle's stable snapshot.
static void device_output(brook::Context &ctx, brook_skeletons &r) {
  auto copy = [&](const void *src, size_t bytes) {
    auto m = std::make_shared<DeviceMemory>(ctx.device(), bytes);
    if (bytes) BROOK_HIP(hipMemcpyAsync(m->ptr, src, bytes, hipMemcpyHostToDevice, ctx.stream()));
    return m;
  };
  r.vertices = copy(r.packed.vertices.data(), r.packed.vertices.size() * sizeof(float));
  r.edges = copy(r.packed.edges.data(), r.packed.edges.size() * sizeof(uint32_t));
  r.radii = copy(r.packed.radii.data(), r.packed.radii.size() * sizeof(float));
  ctx.synchronize();
}

// Rows x width elements, row major, as (rows, width, 1) with byte strides (4 * width, 4, bytes)
// (the reference's row_view, packed.cpp:7-14); width 1: (rows, 1, 1), strides (4, bytes, bytes).
static brook_volume device_rows(const std::shared_ptr<DeviceMemory> &m, brook_dtype dtype, int width, int64_t rows) {
  return view_of_array(rows_array(m, dtype, width, rows));
}
}  // namespace

brook_status brook_skeletonize(brook_context *ctx, const brook_volume *labels, const brook_skeletonize_options *options,
                               brook_skeletons **out) {
  return protect([&] {
    require(ctx && labels && options && out);
    *out = nullptr;
    brook::ContextCall call(ctx->value);
    auto r = std::make_unique<brook_skeletons>();
    r->packed = run_one(ctx->value, labels, options);
    if (options->flags & BROOK_DEVICE_OUTPUT) device_output(ctx->value, *r);
    *out = r.release();
  });
}

namespace {
static brook_status batch(brook_context *ctx, const brook_volume *const *volumes, const brook_volume *const *graphs,
                   size_t count, const brook_skeletonize_options *options, brook_skeletons **out) {
  return protect([&] {
    require(ctx && (volumes || !count) && options && (out || !count));
    for (size_t i = 0; i < count; ++i) out[i] = nullptr;
    require(!options->voxel_graph, "a batch takes its voxel graphs per sample (brook_skeletonize_batch_graphs)");
    brook::ContextCall call(ctx->value);
    std::vector<uint64_t> ids;
    brook::SkeletonizeOptions o = convert(options, &ids);
    std::vector<Input> inputs(count);
    std::vector<GraphInput> graph_inputs(count);
    std::vector<brook::Buffer<uint8_t>> masked(count);
    std::vector<brook::LabelVolume> views;
    for (size_t i = 0; i < count; ++i) {
      views.push_back(engine_input(ctx->value, volumes[i], options, ids, inputs[i], masked[i]));
      if (graphs && graphs[i]) {
        graph_inputs[i] = normalize_graph(ctx->value, graphs[i], volumes[i]);
        views.back().graph = graph_inputs[i].words;
        views.back().graph_compact_2d = graph_inputs[i].compact_2d;
      }
    }
    std::vector<brook::Packed> packed = brook::skeletonize_batch(ctx->value, views, o);
    std::vector<std::unique_ptr<brook_skeletons>> results;
    for (size_t i = 0; i < count; ++i) {
      auto r = std::make_unique<brook_skeletons>();
      r->packed = std::move(packed[i]);
      if (options->flags & BROOK_DEVICE_OUTPUT) device_output(ctx->value, *r);
      results.push_back(std::move(r));
    }
    for (size_t i = 0; i < count; ++i) out[i] = results[i].release();
  });
}
}  // namespace

brook_status brook_skeletonize_batch(brook_context *ctx, const brook_volume *const *volumes, size_t count,
                                     const brook_skeletonize_options *options, brook_skeletons **out) {
  return batch(ctx, volumes, nullptr, count, options, out);
}

brook_status brook_skeletonize_batch_graphs(brook_context *ctx, const brook_volume *const *volumes,
                                            const brook_volume *const *graphs, size_t count,
                                            const brook_skeletonize_options *options, brook_skeletons **out) {
  return batch(ctx, volumes, graphs, count, options, out);
}

brook_status brook_skeletons_get_view(const brook_skeletons *r, brook_skeletons_view *out) {
  return protect([&] {
    require(r && out);
    *out = brook_skeletons_view{};
    const brook::Packed &p = r->packed;
    out->skeleton_count = p.labels.size();
    out->vertex_count = p.radii.size();
    out->edge_count = p.edges.size() / 2;
    out->labels = p.labels.data();
    out->vertex_offsets = reinterpret_cast<const int64_t *>(p.v_off.data());
    out->edge_offsets = reinterpret_cast<const int64_t *>(p.e_off.data());
    out->vertices = p.vertices.data();
    out->edges = p.edges.data();
    out->radii = p.radii.data();
    std::copy(p.anisotropy.begin(), p.anisotropy.end(), out->anisotropy);
  });
}
brook_status brook_skeletons_get_device_view(const brook_skeletons *r, brook_skeletons_device_view *out) {
  return protect([&] {
    require(r && out);
    *out = brook_skeletons_device_view{};
    if (!r->vertices) throw std::invalid_argument("skeleton result was not requested with BROOK_DEVICE_OUTPUT");
    const brook::Packed &p = r->packed;
    out->skeleton_count = p.labels.size();
    out->vertex_count = p.radii.size();
    out->edge_count = p.edges.size() / 2;
    out->labels = p.labels.data();
    out->vertex_offsets = reinterpret_cast<const int64_t *>(p.v_off.data());
    out->edge_offsets = reinterpret_cast<const int64_t *>(p.e_off.data());
    out->vertices = device_rows(r->vertices, BROOK_F32, 3, int64_t(out->vertex_count));
    out->edges = device_rows(r->edges, BROOK_U32, 2, int64_t(out->edge_count));
    out->radii = device_rows(r->radii, BROOK_F32, 1, int64_t(out->vertex_count));
    std::copy(p.anisotropy.begin(), p.anisotropy.end(), out->anisotropy);
  });
}

brook_status brook_skeletons_get_array(const brook_skeletons *r, brook_skeleton_buffer buffer, brook_array **out) {
  return protect([&] {
    require(r && out);
    *out = nullptr;
    if (!r->vertices) throw std::invalid_argument("skeleton result was not requested with BROOK_DEVICE_OUTPUT");
    const int64_t nv = int64_t(r->packed.radii.size()), ne = int64_t(r->packed.edges.size() / 2);
    // the handle's storage, shared: the array stays valid after the handle is destroyed
    switch (buffer) {
      case BROOK_VERTICES: *out = new brook_array(rows_array(r->vertices, BROOK_F32, 3, nv)); break;
      case BROOK_EDGES: *out = new brook_array(rows_array(r->edges, BROOK_U32, 2, ne)); break;
      case BROOK_RADII: *out = new brook_array(rows_array(r->radii, BROOK_F32, 1, nv)); break;
      default: throw std::invalid_argument("invalid skeleton buffer");
    }
  });
}

void brook_skeletons_destroy(brook_skeletons *r) { delete r; }

namespace {
static brook::PackedView view_of(const brook_skeletons_view &v) {
  require(v.vertex_offsets && v.edge_offsets && (v.labels || !v.skeleton_count));
  require((v.vertices && v.radii) || !v.vertex_count);
  require(v.edges || !v.edge_count);
  require(v.vertex_offsets[v.skeleton_count] == int64_t(v.vertex_count) &&
          v.edge_offsets[v.skeleton_count] == int64_t(v.edge_count));
  brook::PackedView p;
  p.skeletons = v.skeleton_count;
  p.labels = v.labels;
  p.v_off = v.vertex_offsets;
  p.e_off = v.edge_offsets;
  p.vertices = v.vertices;
  p.edges = v.edges;
  p.radii = v.radii;
  for (int a = 0; a < 3; ++a) p.anisotropy[a] = v.anisotropy[a];
  return p;
}
static brook_skeletons *merged(brook::Context &ctx, const std::vector<brook::PackedView> &parts, const float *origins) {
  std::vector<std::array<float, 3>> at(parts.size(), std::array<float, 3>{0, 0, 0});
  if (origins)
    for (size_t i = 0; i < parts.size(); ++i) at[i] = {origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]};
  auto r = std::make_unique<brook_skeletons>();
  r->packed = brook::merge_packed(ctx, parts, at);
  device_output(ctx, *r);  // merges return device-output handles (brook.h:234)
  return r.release();
}
}  // namespace

brook_status brook_merge_skeletons(brook_context *ctx, const brook_skeletons *const *fragments, const float *origins,
                                   size_t count, brook_skeletons **out) {
  return protect([&] {
    require(ctx && out && (fragments || !count));
    *out = nullptr;
    brook::ContextCall call(ctx->value);
    std::vector<brook::PackedView> parts;
    std::vector<brook_skeletons_view> views(count);
    for (size_t i = 0; i < count; ++i) {
      require(fragments[i]);
      const brook::Packed &p = fragments[i]->packed;
      brook_skeletons_view &v = views[i];
      v.skeleton_count = p.labels.size();
      v.vertex_count = p.radii.size();
      v.edge_count = p.edges.size() / 2;
      v.labels = p.labels.data();
      v.vertex_offsets = reinterpret_cast<const int64_t *>(p.v_off.data());
      v.edge_offsets = reinterpret_cast<const int64_t *>(p.e_off.data());
      v.vertices = p.vertices.data();
      v.edges = p.edges.data();
      v.radii = p.radii.data();
      std::copy(p.anisotropy.begin(), p.anisotropy.end(), v.anisotropy);
      parts.push_back(view_of(v));
    }
    *out = merged(ctx->value, parts, origins);
  });
}

namespace {
static brook_skeletons *postprocessed(brook::Context &ctx, const brook::PackedView &view, int step, double a, double b) {
  require(step >= BROOK_POST_ALL && step <= BROOK_POST_PRUNE, "unknown postprocess step");
  brook::DevicePacked in = brook::upload_packed(ctx, view);
  brook::DevicePacked done = brook::postprocess(ctx, std::move(in), brook::PostStep(step), a, b);
  auto r = std::make_unique<brook_skeletons>();
  r->packed = brook::download_packed(ctx, done);
  device_output(ctx, *r);  // as merges: a device-output handle (brook.h:234-237)
  return r.release();
}
static brook_skeletons_view view_of_result(const brook::Packed &p) {
  brook_skeletons_view v{};
  v.skeleton_count = p.labels.size();
  v.vertex_count = p.radii.size();
  v.edge_count = p.edges.size() / 2;
  v.labels = p.labels.data();
  v.vertex_offsets = reinterpret_cast<const int64_t *>(p.v_off.data());
  v.edge_offsets = reinterpret_cast<const int64_t *>(p.e_off.data());
  v.vertices = p.vertices.data();
  v.edges = p.edges.data();
  v.radii = p.radii.data();
  std::copy(p.anisotropy.begin(), p.anisotropy.end(), v.anisotropy);
  return v;
}
}  // namespace

brook_status brook_postprocess_canonical(brook_context *ctx, const brook_skeletons *input, double dust, double ticks,
                                         brook_skeletons **out) {
  return protect([&] {
    require(ctx && input && out);
    *out = nullptr;
    brook::ContextCall call(ctx->value);
    const brook_skeletons_view v = view_of_result(input->packed);
    *out = postprocessed(ctx->value, view_of(v), BROOK_POST_ALL, dust, ticks);
  });
}

brook_status brook_postprocess_views(brook_context *ctx, const brook_skeletons_view *input, int step, double a, double b,
                                     brook_skeletons **out) {
  return protect([&] {
    require(ctx && input && out);
    *out = nullptr;
    brook::ContextCall call(ctx->value);
    *out = postprocessed(ctx->value, view_of(*input), step, a, b);
  });
}

brook_status brook_skeleton_components(brook_context *ctx, const brook_skeletons_view *input, int64_t *out, size_t capacity) {
  return protect([&] {
    require(ctx && input && (out || !input->vertex_count));
    require(capacity >= input->vertex_count, "output buffer is too small");
    brook::ContextCall call(ctx->value);
    brook::DevicePacked p = brook::upload_packed(ctx->value, view_of(*input));
    const std::vector<int64_t> roots = brook::skeleton_components(ctx->value, p);
    std::copy(roots.begin(), roots.end(), out);
  });
}

brook_status brook_merge_views(brook_context *ctx, const brook_skeletons_view *fragments, const float *origins, size_t count,
                               brook_skeletons **out) {
  return protect([&] {
    require(ctx && out && (fragments || !count));
    *out = nullptr;
    brook::ContextCall call(ctx->value);
    std::vector<brook::PackedView> parts;
    for (size_t i = 0; i < count; ++i) parts.push_back(view_of(fragments[i]));
    *out = merged(ctx->value, parts, origins);
  });
}

brook_status brook_prepare_fields(brook_context *ctx, const brook_volume *labels, const brook_skeletonize_options *options,
                                  float *dbf, float *daf, float *pdrf, size_t cap, uint8_t *traced, int64_t *root,
                                  int64_t *target, float *dbf_max, float *max_daf, int64_t *count) {
  return protect([&] {
    require(ctx && labels && options);
    brook::ContextCall call(ctx->value);
    brook::FieldsOut f;
    f.dbf = dbf;
    f.daf = daf;
    f.pdrf = pdrf;
    f.stop = true;
    run_one(ctx->value, labels, options, &f);
    require(f.active.size() <= cap || (!traced && !root && !target && !dbf_max && !max_daf), "table capacity too small");
    for (size_t i = 0; i < f.active.size(); ++i) {
      if (traced) traced[i] = f.active[i];
      if (root) root[i] = f.root[i];
      if (target) target[i] = f.target[i];
      if (dbf_max) dbf_max[i] = f.dbf_max[i];
      if (max_daf) max_daf[i] = f.max_daf[i];
    }
    if (count) *count = f.count;
  });
}

brook_status brook_target_order(brook_context *ctx, const brook_volume *ids, const float *daf, int64_t *order,
                                size_t capacity, size_t *count) {
  return protect([&] {
    require(ctx && ids && count && (daf || !ids->shape[0]));
    *count = 0;
    require(ids->dtype == BROOK_U8 || ids->dtype == BROOK_U16 || ids->dtype == BROOK_U32,
            "target orders take unsigned component ids of at most 32 bits");
    brook::ContextCall call(ctx->value);
    Input in = normalize(ctx->value, ids);
    const std::vector<int> got = brook::target_orders(ctx->value, in.view, daf);
    *count = got.size();
    require(got.size() <= capacity && (order || got.empty()), "order capacity below the foreground voxel count");
    std::copy(got.begin(), got.end(), order);
  });
}

brook_status brook_context_stat(const brook_context *ctx, const char *name, double *value) {
  return protect([&] {
    require(ctx && name && value);
    auto it = ctx->value.stats.find(name);
    *value = it == ctx->value.stats.end() ? 0.0 : it->second;
  });
}

brook_status brook_context_wait_stream(brook_context *ctx, void *stream) {
  return protect([&] {
    require(ctx != nullptr);
    brook::ContextCall call(ctx->value);
    hipEvent_t event = nullptr;
    BROOK_HIP(hipEventCreateWithFlags(&event, hipEventDisableTiming));
    hipError_t e = hipEventRecord(event, static_cast<hipStream_t>(stream));
    if (e == hipSuccess) e = hipStreamWaitEvent(ctx->value.stream(), event, 0);
    (void)hipEventDestroy(event);
    BROOK_HIP(e);
  });
}

brook_status brook_context_set_kernel_timing(brook_context *ctx, int enable) {
  return protect([&] {
    require(ctx != nullptr);
    ctx->value.kernel_timing = enable != 0;
  });
}

}  // extern "C"
