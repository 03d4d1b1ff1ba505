// C++17 RAII wrapper over the C ABI (the reference's cpp/include/brook/brook.hpp:29-194: the
// brook::api namespace with Context, Array, Components and Skeletons). Handles are move-only and
// destroy their C handle; every failing call throws std::runtime_error with brook_last_error().
#pragma once
#include <array>
#include <memory>
#include <stdexcept>
#include <utility>
#include <vector>

#include "brook.h"

namespace brook {
namespace api {

inline void check(brook_status status) {
  if (status != BROOK_SUCCESS) throw std::runtime_error(brook_last_error());
}

// The Python skeletonize() defaults (brook_skeletonize_options_init).
inline brook_skeletonize_options skeletonize_defaults() {
  brook_skeletonize_options options;
  brook_skeletonize_options_init(&options);
  return options;
}

namespace detail {
template <class T, void (*Destroy)(T *)> struct Destroyer {
  void operator()(T *p) const noexcept {
    if (p) Destroy(p);
  }
};
template <class T, void (*Destroy)(T *)> using Owned = std::unique_ptr<T, Destroyer<T, Destroy>>;
inline void destroy_tensor(DLManagedTensor *t) {
  if (t->deleter) t->deleter(t);
}
inline void destroy_versioned(DLManagedTensorVersioned *t) {
  if (t->deleter) t->deleter(t);
}
}  // namespace detail

// A DLPack export that calls the producer's deleter once, when it goes out of scope.
using DLPackTensor = detail::Owned<DLManagedTensor, detail::destroy_tensor>;
using DLPackTensorVersioned = detail::Owned<DLManagedTensorVersioned, detail::destroy_versioned>;

// A device array (a volume or a skeleton buffer's rows).
class Array {
 public:
  explicit Array(brook_array *handle) : handle_(handle) {}
  brook_array *get() const noexcept { return handle_.get(); }
  brook_volume view() const {
    brook_volume out{};
    check(brook_array_view(handle_.get(), &out));
    return out;
  }
  void copy_to_host(void *destination, size_t bytes) const {
    check(brook_array_copy_to_host(handle_.get(), destination, bytes));
  }
  // Zero copy: the tensor keeps the storage alive after this Array and its context are gone.
  DLPackTensor dlpack() const {
    DLManagedTensor *out = nullptr;
    check(brook_array_to_dlpack(handle_.get(), &out));
    return DLPackTensor(out);
  }
  DLPackTensorVersioned dlpack_versioned(uint32_t max_minor = DLPACK_MINOR_VERSION) const {
    DLManagedTensorVersioned *out = nullptr;
    check(brook_array_to_dlpack_versioned(handle_.get(), max_minor, &out));
    return DLPackTensorVersioned(out);
  }

 private:
  detail::Owned<brook_array, brook_array_destroy> handle_;
};

struct ComponentsView {
  brook_volume labels{};
  const int64_t *mapping = nullptr, *roots = nullptr;
  size_t count = 0;
};

// Connected components: device ids plus host mapping and roots.
class Components {
 public:
  explicit Components(brook_components *handle) : handle_(handle) {}
  brook_components *get() const noexcept { return handle_.get(); }
  ComponentsView view() const {
    ComponentsView out;
    check(brook_components_view(handle_.get(), &out.labels, &out.mapping, &out.roots, &out.count));
    return out;
  }
  void copy_to_host(void *destination, size_t bytes) const {
    check(brook_components_copy_to_host(handle_.get(), destination, bytes));
  }
  // The ids as an Array that shares the storage (valid after this handle is destroyed).
  Array labels() const {
    brook_array *out = nullptr;
    check(brook_components_labels(handle_.get(), &out));
    return Array(out);
  }

 private:
  detail::Owned<brook_components, brook_components_destroy> handle_;
};

// A packed skeleton result: host view, and (BROOK_DEVICE_OUTPUT) device view and arrays.
class Skeletons {
 public:
  explicit Skeletons(brook_skeletons *handle) : handle_(handle) {}
  brook_skeletons *get() const noexcept { return handle_.get(); }
  brook_skeletons_view view() const {
    brook_skeletons_view out{};
    check(brook_skeletons_get_view(handle_.get(), &out));
    return out;
  }
  brook_skeletons_device_view device_view() const {
    brook_skeletons_device_view out{};
    check(brook_skeletons_get_device_view(handle_.get(), &out));
    return out;
  }
  Array array(brook_skeleton_buffer buffer) const {
    brook_array *out = nullptr;
    check(brook_skeletons_get_array(handle_.get(), buffer, &out));
    return Array(out);
  }

 private:
  detail::Owned<brook_skeletons, brook_skeletons_destroy> handle_;
};

// A context on one device; calls block and are serialized per context.
class Context {
 public:
  explicit Context(int device = 0) {
    brook_context *p = nullptr;
    check(brook_context_create(device, &p));
    handle_.reset(p);
  }
  brook_context *get() const noexcept { return handle_.get(); }
  void synchronize() const { check(brook_context_synchronize(handle_.get())); }
  // EXTENSION: later work on this context waits for the work queued so far on `stream`.
  void wait_stream(void *stream) const { check(brook_context_wait_stream(handle_.get(), stream)); }
  Array upload(const brook_volume &volume) {
    brook_array *out = nullptr;
    check(brook_array_upload(handle_.get(), &volume, &out));
    return Array(out);
  }
  Components connected_components(const brook_volume &volume) {
    brook_components *out = nullptr;
    check(brook_connected_components(handle_.get(), &volume, &out));
    return Components(out);
  }
  Array edt(const brook_volume &volume, const float anisotropy[3], bool black_border = false) {
    brook_array *out = nullptr;
    check(brook_edt(handle_.get(), &volume, anisotropy, black_border ? 1 : 0, &out));
    return Array(out);
  }
  Skeletons skeletonize(const brook_volume &volume, const brook_skeletonize_options &options = skeletonize_defaults()) {
    brook_skeletons *out = nullptr;
    check(brook_skeletonize(handle_.get(), &volume, &options, &out));
    return Skeletons(out);
  }
  // EXTENSION: independent volumes in one call; result i equals skeletonize(volumes[i]).
  std::vector<Skeletons> skeletonize_batch(const std::vector<brook_volume> &volumes,
                                           const brook_skeletonize_options &options = skeletonize_defaults()) {
    std::vector<const brook_volume *> pointers;
    for (const brook_volume &v : volumes) pointers.push_back(&v);
    std::vector<brook_skeletons *> handles(volumes.size(), nullptr);
    check(brook_skeletonize_batch(handle_.get(), pointers.data(), pointers.size(), &options, handles.data()));
    std::vector<Skeletons> out;
    for (brook_skeletons *h : handles) out.emplace_back(h);
    return out;
  }
  // Fragments offset by their origins merged into one result (a device-output handle).
  Skeletons merge(const std::vector<const Skeletons *> &fragments, const std::vector<std::array<float, 3>> &origins) {
    if (fragments.size() != origins.size()) throw std::invalid_argument("fragment/origin count mismatch");
    std::vector<const brook_skeletons *> handles;
    std::vector<float> flat;
    for (size_t i = 0; i < fragments.size(); ++i) {
      handles.push_back(fragments[i] ? fragments[i]->get() : nullptr);
      flat.insert(flat.end(), origins[i].begin(), origins[i].end());
    }
    brook_skeletons *out = nullptr;
    check(brook_merge_skeletons(handle_.get(), handles.data(), flat.empty() ? nullptr : flat.data(), handles.size(), &out));
    return Skeletons(out);
  }
  // The canonical postprocess on the device (brook.hpp:177-178 of the reference): dust, loops,
  // joins, ticks; a device-output handle.
  Skeletons postprocess_canonical(const Skeletons &input, double dust_threshold = 1500, double tick_threshold = 3500) {
    brook_skeletons *out = nullptr;
    check(brook_postprocess_canonical(handle_.get(), input.get(), dust_threshold, tick_threshold, &out));
    return Skeletons(out);
  }

 private:
  detail::Owned<brook_context, brook_context_destroy> handle_;
};

}  // namespace api
}  // namespace brook
