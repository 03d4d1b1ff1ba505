// A C++ consumer of include/brook.hpp, built against libbrook_amd.so by __graft_entry__.build() and
// run by tests/test_consumers.py on a GPU, modelled on the reference's C++ test
// (cpp/tests/cpp_smoke.cpp): a 28-voxel bar of label 42 through the RAII handles, device output,
// DLPack views of skeleton buffers, merge, postprocess and batch. Exit status 0 = pass; else the failed check.
#include <cstdio>
#include <cstring>
#include <vector>

#include "brook.hpp"

#define REQUIRE(cond, code)                         \
  do {                                              \
    if (!(cond)) {                                  \
      std::fprintf(stderr, "check %d failed\n", code); \
      return code;                                  \
    }                                               \
  } while (0)

int main() try {
  std::vector<uint32_t> image(32 * 8 * 8, 0);
  for (int z = 3; z < 5; ++z)
    for (int y = 3; y < 5; ++y)
      for (int x = 2; x < 30; ++x) image[x + 32 * (y + 8 * z)] = 42;
  brook_volume volume{};
  volume.struct_size = sizeof volume;
  volume.abi_version = BROOK_ABI_VERSION;
  volume.data = image.data();
  volume.dtype = BROOK_U32;
  volume.memory = BROOK_HOST;
  volume.shape[0] = 32;
  volume.shape[1] = volume.shape[2] = 8;
  volume.strides[0] = 4;
  volume.strides[1] = 128;
  volume.strides[2] = 1024;
  brook_skeletonize_options options = brook::api::skeletonize_defaults();
  options.dust_threshold = 0;
  options.teasar.scale = 1.5;
  options.teasar.constant = 2;
  options.flags = BROOK_FIX_BRANCHING;

  // the component ids outlive their context and components handle
  brook::api::Array ids = [&] {
    brook::api::Context context;
    brook::api::Components components = context.connected_components(volume);
    brook::api::ComponentsView view = components.view();
    if (view.count != 1 || view.mapping[1] != 42 || view.roots[0] != 2 + 32 * (3 + 8 * 3))
      throw std::runtime_error("wrong component mapping");
    return components.labels();
  }();
  brook::api::DLPackTensor id_tensor = ids.dlpack();
  REQUIRE(id_tensor->dl_tensor.shape[0] == 32 && id_tensor->dl_tensor.device.device_type == kDLROCM, 1);
  std::vector<uint8_t> host_ids(32 * 8 * 8);
  ids.copy_to_host(host_ids.data(), host_ids.size());
  for (size_t i = 0; i < image.size(); ++i) REQUIRE(host_ids[i] == (image[i] ? 1 : 0), 2);

  brook::api::Context context;
  brook::api::Skeletons result = context.skeletonize(volume, options);
  const brook_skeletons_view data = result.view();
  REQUIRE(data.skeleton_count == 1 && data.labels[0] == 42 && data.edge_count && data.vertex_count, 3);
  REQUIRE(data.vertex_offsets[1] == int64_t(data.vertex_count) && data.edge_offsets[1] == int64_t(data.edge_count), 4);

  options.flags |= BROOK_DEVICE_OUTPUT;
  brook::api::Skeletons device_result = context.skeletonize(volume, options);
  const brook_skeletons_device_view dv = device_result.device_view();
  REQUIRE(dv.vertex_count == data.vertex_count && dv.edge_count == data.edge_count, 5);
  REQUIRE(dv.vertices.memory == BROOK_DEVICE && dv.vertices.shape[0] == int64_t(data.vertex_count) &&
              dv.vertices.shape[1] == 3 && dv.vertices.strides[0] == 12 && dv.vertices.strides[1] == 4,
          6);
  REQUIRE(dv.edges.shape[1] == 2 && dv.edges.strides[0] == 8 && dv.radii.shape[0] == int64_t(data.vertex_count) &&
              dv.radii.shape[1] == 1 && dv.radii.strides[0] == 4,
          7);
  brook::api::Array vertices = device_result.array(BROOK_VERTICES);
  const brook_volume av = vertices.view();
  REQUIRE(av.shape[0] == dv.vertices.shape[0] && av.shape[1] == 3 && av.strides[0] == 12, 8);
  brook::api::DLPackTensor vt = vertices.dlpack();
  REQUIRE(vt->dl_tensor.ndim == 2 && vt->dl_tensor.shape[1] == 3 && vt->dl_tensor.strides[0] == 3 &&
              vt->dl_tensor.strides[1] == 1,
          9);
  brook::api::DLPackTensor rt = device_result.array(BROOK_RADII).dlpack();
  REQUIRE(rt->dl_tensor.ndim == 1 && rt->dl_tensor.shape[0] == int64_t(data.vertex_count), 10);
  std::vector<float> device_vertices(3 * data.vertex_count);
  vertices.copy_to_host(device_vertices.data(), device_vertices.size() * sizeof(float));
  REQUIRE(std::memcmp(device_vertices.data(), data.vertices, device_vertices.size() * sizeof(float)) == 0, 11);
  const brook_skeletons_view snapshot = device_result.view();
  for (size_t i = 0; i < 2 * data.edge_count; ++i) REQUIRE(snapshot.edges[i] == data.edges[i], 12);

  // two identical fragments at the same origin merge into the skeleton itself
  brook::api::Skeletons merged = context.merge({&result, &device_result}, {{{0, 0, 0}}, {{0, 0, 0}}});
  const brook_skeletons_view after = merged.view();
  REQUIRE(after.skeleton_count == 1 && after.vertex_count == data.vertex_count && after.edge_count == data.edge_count, 13);
  for (size_t i = 0; i < 3 * data.vertex_count; ++i) REQUIRE(after.vertices[i] == data.vertices[i], 14);
  for (size_t i = 0; i < 2 * data.edge_count; ++i) REQUIRE(after.edges[i] == data.edges[i], 15);
  // the canonical postprocess without dust or tick thresholds leaves it as it is (cpp_smoke.cpp:44-48)
  brook::api::Skeletons processed = context.postprocess_canonical(merged, 0, 0);
  const brook_skeletons_view post = processed.view();
  REQUIRE(post.skeleton_count == 1 && post.vertex_count == data.vertex_count && post.edge_count == data.edge_count, 19);
  for (size_t i = 0; i < 3 * data.vertex_count; ++i) REQUIRE(post.vertices[i] == data.vertices[i], 20);
  for (size_t i = 0; i < 2 * data.edge_count; ++i) REQUIRE(post.edges[i] == data.edges[i], 21);

  // a batch of the volume and its mirror image equals separate calls
  std::vector<uint32_t> mirror(image.size());
  for (int z = 0; z < 8; ++z)
    for (int y = 0; y < 8; ++y)
      for (int x = 0; x < 32; ++x) mirror[x + 32 * (y + 8 * z)] = image[(31 - x) + 32 * (y + 8 * z)];
  brook_volume second = volume;
  second.data = mirror.data();
  options.flags = BROOK_FIX_BRANCHING;
  std::vector<brook::api::Skeletons> batch = context.skeletonize_batch({volume, second}, options);
  brook::api::Skeletons alone = context.skeletonize(second, options);
  const brook_skeletons_view b0 = batch[0].view(), b1 = batch[1].view(), a1 = alone.view();
  REQUIRE(b0.vertex_count == data.vertex_count && b1.vertex_count == a1.vertex_count && b1.edge_count == a1.edge_count, 16);
  REQUIRE(std::memcmp(b1.vertices, a1.vertices, 12 * a1.vertex_count) == 0, 17);

  // errors surface as exceptions carrying brook_last_error()
  bool threw = false;
  try {
    brook::api::Context bad(-1);
  } catch (const std::runtime_error &) {
    threw = true;
  }
  REQUIRE(threw, 18);
  std::printf("cpp consumer ok\n");
  return 0;
} catch (const std::exception &e) {
  std::fprintf(stderr, "exception: %s\n", e.what());
  return 100;
}
