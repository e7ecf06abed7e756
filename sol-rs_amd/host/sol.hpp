// sol.hpp -- C++ mirror of the sol-rs ray-tracing API over the C ABI of
// include/solhip.h.
//
// Same names, argument meaning and ownership as the reference's Rust API:
//   sol::scene::{ModelVertex, MaterialInfo, BufferPart, PrimitiveSection,
//                Mesh, Scene, Camera, load_scene}      (src/scene/*.rs)
//   sol::ray::{SceneInstance, GeometryInstance, BLAS, TLAS,
//              SceneDescription, PipelineInfo, Pipeline,
//              ShaderBindingTableInfo, ShaderBindingTable} (src/ray/*.rs)
// Differences forced by the hardware: a CommandBuffer records the bindings
// (pipeline, scene, images, uniforms, push constant) that Vulkan descriptor
// sets carry, and is submitted to a HIP stream; errors throw sol::Error where
// the reference panics.
#pragma once

#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/solhip.h"
#include "sol_math.hpp"

namespace sol {

struct Error : std::runtime_error {
    int32_t code;
    Error(int32_t c, const std::string &m) : std::runtime_error(m), code(c) {}
};
void check(int32_t rc, const char *what);

// device + default stream (SharedContext, src/context.rs:239-370)
class Context {
  public:
    explicit Context(int device = 0);
    ~Context();
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    solhip_ctx handle() const { return ctx_; }
    solhip_stream stream() const { return 0; }
    void wait_idle() const;

  private:
    solhip_ctx ctx_ = 0;
};

// device buffer (src/buffer.rs): from_data uploads and blocks, update writes
class Buffer {
  public:
    Buffer() = default;
    Buffer(std::shared_ptr<Context> ctx, uint64_t bytes);
    template <class T>
    static Buffer from_data(std::shared_ptr<Context> ctx, const std::vector<T> &data)
    {
        Buffer b(std::move(ctx), data.size() * sizeof(T));
        b.upload(data.data(), data.size() * sizeof(T), 0);
        b.count_ = data.size();
        return b;
    }
    ~Buffer();
    Buffer(Buffer &&o) noexcept;
    Buffer &operator=(Buffer &&o) noexcept;
    Buffer(const Buffer &) = delete;
    Buffer &operator=(const Buffer &) = delete;

    void upload(const void *src, uint64_t bytes, uint64_t offset);
    template <class T>
    void update(const std::vector<T> &data)
    {
        upload(data.data(), data.size() * sizeof(T), 0);
    }
    void download(void *dst, uint64_t bytes, uint64_t offset = 0) const;
    uint64_t get_device_address() const { return addr_; }
    uint64_t get_size() const { return bytes_; }
    uint64_t get_element_count() const { return count_; }
    solhip_buffer handle() const { return buf_; }

  private:
    std::shared_ptr<Context> ctx_;
    solhip_buffer buf_ = 0;
    uint64_t addr_ = 0, bytes_ = 0, count_ = 0;
};

// byte range of a buffer (vk::DescriptorBufferInfo)
struct DescriptorBufferInfo {
    uint64_t address = 0;
    uint64_t offset = 0;
    uint64_t range = 0;
};

namespace scene {

// src/scene/mesh.rs:7-25
struct ModelVertex {
    Vec4 pos{0, 0, 0, 1};
    Vec4 color{1, 1, 1, 1};
    Vec4 normal{0, 0, 0, 0};
    Vec4 uv{0, 0, 0, 0};
    static uint32_t stride() { return sizeof(ModelVertex); }
};
static_assert(sizeof(ModelVertex) == 64, "ModelVertex");

// src/scene/mod.rs:19-29
struct MaterialInfo {
    Vec4 base_color{0, 0, 0, 0};
    Vec3 emissive_factor{0, 0, 0};
    float padding0 = 0;
    float metallic_factor = 0, roughness_factor = 0, padding1 = 0, padding2 = 0;
};
static_assert(sizeof(MaterialInfo) == 48, "MaterialInfo");

struct BufferPart {
    size_t offset = 0, element_count = 0;
};

// src/scene/mod.rs:37-97
class PrimitiveSection {
  public:
    size_t index = 0;
    BufferPart vertices;
    std::optional<BufferPart> indices;
    std::optional<size_t> material_index;

    template <class T>
    DescriptorBufferInfo get_index_descriptor(const Buffer &b) const
    {
        return {b.get_device_address(), indices.value().offset * sizeof(T), indices.value().element_count * sizeof(T)};
    }
    DescriptorBufferInfo get_vertex_descriptor(const Buffer &b) const
    {
        return {b.get_device_address(), vertices.offset * sizeof(ModelVertex),
                vertices.element_count * sizeof(ModelVertex)};
    }
    DescriptorBufferInfo get_material_descriptor(const Buffer &b) const
    {
        return {b.get_device_address(), material_index.value() * sizeof(MaterialInfo), sizeof(MaterialInfo)};
    }
    uint32_t get_vertex_count() const { return (uint32_t)vertices.element_count; }
    uint32_t get_vertex_offset() const { return (uint32_t)vertices.offset; }
    uint64_t get_vertex_offset_size() const { return vertices.offset * sizeof(ModelVertex); }
    uint32_t get_index_count() const { return (uint32_t)indices.value().element_count; }
    template <class T>
    uint64_t get_index_offset_size() const
    {
        return indices.value().offset * sizeof(T);
    }
};

struct Mesh {
    std::string name;
    Buffer vertex_buffer;
    std::optional<Buffer> index_buffer;    // u32
    std::optional<Buffer> index_storage;   // u64 (what the hit shaders read)
    Mat4 transform;
    std::vector<PrimitiveSection> primitive_sections;
    // host copies (for dumps / tests)
    std::vector<ModelVertex> host_vertices;
    std::vector<uint32_t> host_indices;
};

// src/scene/camera.rs:54-113 (mouse manipulation is not part of the ray path)
class Camera {
  public:
    explicit Camera(Vec2 window_size);
    static Camera from_view(Mat4 view, float yfov, float z_near, float z_far);
    void look_at(Vec3 eye, Vec3 center, Vec3 up);
    void set_window_size(Vec2 window_size);
    Mat4 view_matrix() const { return view_matrix_; }
    Mat4 perspective_matrix() const { return persp_matrix_; }

  private:
    Camera() = default;
    void update_view();
    void update_persp();
    Vec3 position_, center_, up_;
    float vfov_ = 35.0f, z_near_ = 0.1f, z_far_ = 1000.0f;
    Mat4 view_matrix_, persp_matrix_;
    Vec2 window_size_;
};

struct Scene {
    std::vector<Mesh> meshes;
    std::vector<MaterialInfo> materials;
    Buffer material_buffer;
    std::optional<Camera> camera;
};

// host side of load_scene: the glTF decoded into meshes (host vertex / index
// arrays, primitive sections, global transforms), materials and camera,
// without touching the device
struct HostMesh {
    std::string name;
    std::vector<ModelVertex> vertices;
    std::vector<uint32_t> indices;
    Mat4 transform;
    std::vector<PrimitiveSection> primitive_sections;
};
struct HostScene {
    std::vector<HostMesh> meshes;
    std::vector<MaterialInfo> materials;
    std::optional<Camera> camera;
};
HostScene decode_scene(const std::string &path);

// glTF -> Scene (src/scene/mod.rs:138-295): decode_scene, then the device
// uploads of Buffer::from_data; throws sol::Error
Scene load_scene(std::shared_ptr<Context> ctx, const std::string &path);

}  // namespace scene

namespace ray {

// src/ray/mod.rs:16-35
struct SceneInstance {
    uint32_t id = 0;
    uint32_t texture_offset = 0;
    Vec2 padding;
    Mat4 transform;
    Mat4 transform_it;
    void update_transform(const Mat4 &t)
    {
        transform = t;
        transform_it = t.inverse().transpose();
    }
    const Mat4 &get_transform() const { return transform; }
};
static_assert(sizeof(SceneInstance) == 144, "SceneInstance");

// src/ray/acceleration.rs:5-14
struct GeometryInstance {
    uint64_t vertex_buffer = 0;
    uint32_t vertex_count = 0;
    uint32_t vertex_offset = 0;
    uint64_t vertex_offset_size = 0;
    std::optional<uint64_t> index_buffer;
    std::optional<uint32_t> index_count;
    std::optional<uint64_t> index_offset_size;
    Mat4 transform;
};

// stands in for vk::CommandBuffer: a stream plus the bindings a trace reads
class Pipeline;
class SceneDescription;
struct Images {
    Buffer *accum = nullptr;       // RGBA32F (binding 1 of both examples)
    Buffer *render = nullptr;      // RGBA8 (pathtrace binding 2)
    Buffer *blue_noise = nullptr;  // RGBA8 texture (AO binding 2)
    uint32_t blue_w = 0, blue_h = 0;
};
struct CommandBuffer {
    solhip_stream stream = 0;
    const Pipeline *pipeline = nullptr;
    const SceneDescription *scene = nullptr;
    Images images;
    solhip_scene_uniforms uniforms{};
    int32_t accum_start_frame = 0;   // push constant
    void bind_pipeline(const Pipeline &p) { pipeline = &p; }
    void bind_scene(const SceneDescription &s) { scene = &s; }
    void bind_images(const Images &i) { images = i; }
    void set_uniforms(const solhip_scene_uniforms &u) { uniforms = u; }
    void push_constants(int32_t accum_start) { accum_start_frame = accum_start; }
};

// src/ray/acceleration.rs:128-248
class BLAS {
  public:
    BLAS(std::shared_ptr<Context> ctx, CommandBuffer &cmd, const std::vector<GeometryInstance> &geometries,
         Mat4 transform, uint64_t vertex_stride, bool is_opaque);
    ~BLAS();
    BLAS(BLAS &&o) noexcept;
    BLAS &operator=(BLAS &&o) noexcept;
    BLAS(const BLAS &) = delete;
    BLAS &operator=(const BLAS &) = delete;
    Mat4 get_transform() const { return transform_; }
    void set_transform(const Mat4 &t) { transform_ = t; }
    solhip_blas handle() const { return blas_; }
    uint32_t hit_group_index() const { return 0; }

  private:
    std::shared_ptr<Context> ctx_;
    solhip_blas blas_ = 0;
    Mat4 transform_;
};

// src/ray/acceleration.rs:308-474
class TLAS {
  public:
    TLAS(std::shared_ptr<Context> ctx, CommandBuffer &cmd, const std::vector<BLAS> &blas);
    ~TLAS();
    TLAS(TLAS &&o) noexcept;
    TLAS &operator=(TLAS &&o) noexcept;
    TLAS(const TLAS &) = delete;
    TLAS &operator=(const TLAS &) = delete;
    void regenerate(CommandBuffer &cmd, const std::vector<BLAS> &blas);
    solhip_tlas handle() const { return tlas_; }
    static std::vector<solhip_instance> create_instances(const std::vector<BLAS> &blas);

  private:
    std::shared_ptr<Context> ctx_;
    solhip_tlas tlas_ = 0;
};

// src/ray/mod.rs:37-206
class SceneDescription {
  public:
    static SceneDescription from_scene(std::shared_ptr<Context> ctx, const scene::Scene &scene);
    static SceneDescription from_meshes(std::shared_ptr<Context> ctx, const std::vector<const scene::Mesh *> &meshes,
                                        const std::vector<Mat4> &mesh_transforms, const Buffer *material_buffer);
    const TLAS &tlas() const { return *tlas_; }
    void blas_transform(const Mat4 &transform, size_t index);
    void blas_transforms(const std::vector<Mat4> &transforms);
    void tlas_regenerate(CommandBuffer &cmd);
    const std::vector<BLAS> &blas() const { return blas_; }
    const Buffer &get_instances_buffer() const { return instances_buffer_; }
    void update();
    const std::vector<DescriptorBufferInfo> &get_vertex_descriptors() const { return vertex_descriptors_; }
    const std::vector<DescriptorBufferInfo> &get_index_descriptors() const { return index_descriptors_; }
    const std::vector<DescriptorBufferInfo> &get_material_descriptors() const { return mat_descriptors_; }
    const std::vector<SceneInstance> &instances() const { return instances_; }
    // the descriptor arrays as device address tables for the programs
    const solhip_scene_tables &tables() const { return tables_; }

  private:
    SceneDescription() = default;
    void build_tables(std::shared_ptr<Context> ctx);
    std::vector<BLAS> blas_;
    std::unique_ptr<TLAS> tlas_;
    std::vector<SceneInstance> instances_;
    Buffer instances_buffer_;
    std::vector<DescriptorBufferInfo> vertex_descriptors_, index_descriptors_, mat_descriptors_;
    std::vector<std::pair<size_t, std::vector<size_t>>> blas_to_instances_;
    Buffer vertex_table_, index_table_, material_table_;
    solhip_scene_tables tables_{};
};

// shader stage of a PipelineInfo entry (vk::ShaderStageFlags subset)
enum class ShaderStage { Raygen, Miss, ClosestHit };

// src/ray/pipeline.rs:5-52
struct PipelineInfo {
    std::vector<std::pair<std::string, ShaderStage>> shaders;
    std::string name_;
    std::vector<uint32_t> specialization_data;
    PipelineInfo &shader(const std::string &path, ShaderStage stage)
    {
        shaders.emplace_back(path, stage);
        return *this;
    }
    PipelineInfo &name(const std::string &n)
    {
        name_ = n;
        return *this;
    }
    // specialization constant 0 = ENABLE_SKYLIGHT (pathtrace.rmiss:5)
    PipelineInfo &specialization(uint32_t value, uint32_t constant_id)
    {
        if (constant_id >= specialization_data.size()) specialization_data.resize(constant_id + 1, 0);
        specialization_data[constant_id] = value;
        return *this;
    }
};

// src/ray/pipeline.rs:54-144: the shader set selects the program
class Pipeline {
  public:
    Pipeline(std::shared_ptr<Context> ctx, const PipelineInfo &info, const solhip_spec *override_spec = nullptr);
    ~Pipeline();
    Pipeline(const Pipeline &) = delete;
    Pipeline &operator=(const Pipeline &) = delete;
    solhip_program handle() const { return prog_; }
    uint32_t kind() const { return kind_; }
    const solhip_spec &spec() const { return spec_; }

  private:
    std::shared_ptr<Context> ctx_;
    solhip_program prog_ = 0;
    uint32_t kind_ = 0;
    solhip_spec spec_{};
};

// src/ray/sbt.rs:13-55 (group indices are recorded; no table is needed)
struct ShaderBindingTableInfo {
    std::vector<uint64_t> raygen_indices, miss_indices, hit_group_indices;
    ShaderBindingTableInfo &raygen(uint64_t i)
    {
        raygen_indices.push_back(i);
        return *this;
    }
    ShaderBindingTableInfo &miss(uint64_t i)
    {
        miss_indices.push_back(i);
        return *this;
    }
    ShaderBindingTableInfo &hitgroup(uint64_t i)
    {
        hit_group_indices.push_back(i);
        return *this;
    }
};

struct Extent3D {
    uint32_t width = 0, height = 0, depth = 1;
};

// src/ray/sbt.rs:57-181
class ShaderBindingTable {
  public:
    ShaderBindingTable(std::shared_ptr<Context> ctx, const Pipeline &pipeline, const ShaderBindingTableInfo &info);
    // one W x H launch of the bound pipeline with the recorded bindings
    void cmd_trace_rays(CommandBuffer &cmd, Extent3D extent) const;

  private:
    std::shared_ptr<Context> ctx_;
    const Pipeline *pipeline_;
};

}  // namespace ray
}  // namespace sol
