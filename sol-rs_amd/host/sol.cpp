// sol.cpp -- see sol.hpp.
#include "sol.hpp"

#include <cstring>
#include <functional>

#include "gltf.hpp"

namespace sol {

void check(int32_t rc, const char *what)
{
    if (rc == SOLHIP_OK) return;
    char buf[1024];
    solhip_last_error(buf, sizeof(buf));
    throw Error(rc, std::string(what) + ": " + buf);
}

// ---- Context / Buffer ---------------------------------------------------------
Context::Context(int device) { check(solhip_context_create(device, 0, &ctx_), "solhip_context_create"); }
Context::~Context()
{
    if (ctx_) solhip_context_destroy(ctx_);
}
void Context::wait_idle() const { check(solhip_stream_sync(ctx_, 0), "solhip_stream_sync"); }

Buffer::Buffer(std::shared_ptr<Context> ctx, uint64_t bytes) : ctx_(std::move(ctx)), bytes_(bytes)
{
    check(solhip_buffer_create(ctx_->handle(), bytes ? bytes : 16, &buf_), "solhip_buffer_create");
    check(solhip_buffer_address(ctx_->handle(), buf_, &addr_), "solhip_buffer_address");
}
Buffer::~Buffer()
{
    if (buf_ && ctx_) solhip_buffer_destroy(ctx_->handle(), buf_);
}
Buffer::Buffer(Buffer &&o) noexcept
    : ctx_(std::move(o.ctx_)), buf_(o.buf_), addr_(o.addr_), bytes_(o.bytes_), count_(o.count_)
{
    o.buf_ = 0;
}
Buffer &Buffer::operator=(Buffer &&o) noexcept
{
    if (this != &o) {
        if (buf_ && ctx_) solhip_buffer_destroy(ctx_->handle(), buf_);
        ctx_ = std::move(o.ctx_);
        buf_ = o.buf_;
        addr_ = o.addr_;
        bytes_ = o.bytes_;
        count_ = o.count_;
        o.buf_ = 0;
    }
    return *this;
}
void Buffer::upload(const void *src, uint64_t bytes, uint64_t offset)
{
    if (bytes) check(solhip_buffer_upload(ctx_->handle(), buf_, offset, src, bytes), "Buffer::upload");
}
void Buffer::download(void *dst, uint64_t bytes, uint64_t offset) const
{
    if (bytes) check(solhip_buffer_download(ctx_->handle(), buf_, offset, dst, bytes), "Buffer::download");
}

namespace scene {

// ---- Camera (src/scene/camera.rs:54-113) -------------------------------------
Camera::Camera(Vec2 window_size)
    : position_(Vec3::splat(10.0f)), center_{0, 0, 0}, up_{0, -1, 0}, vfov_(35.0f), z_near_(0.1f), z_far_(1000.0f),
      window_size_(window_size)
{
    update_persp();
}

Camera Camera::from_view(Mat4 view, float yfov, float z_near, float z_far)
{
    const Mat4 vi = view.inverse();
    const Vec4 position = vi * Vec4{0, 0, 0, 1};
    const Vec4 up = vi * Vec4{0, 1, 0, 0};
    const Vec4 fwd = vi * Vec4{0, 0, -4, 0};
    Camera c;
    c.position_ = position.xyz();
    c.center_ = Vec3{position.x + fwd.x, position.y + fwd.y, position.z + fwd.z};
    c.up_ = up.xyz();
    c.vfov_ = yfov;  // used as degrees by update_persp (src/scene/camera.rs:105)
    c.z_near_ = z_near;
    c.z_far_ = z_far;
    c.view_matrix_ = view;
    c.persp_matrix_ = Mat4::identity();
    c.window_size_ = Vec2{1920.0f, 1080.0f};
    return c;
}

void Camera::update_view() { view_matrix_ = Mat4::look_at_rh(position_, center_, up_); }
void Camera::update_persp()
{
    const float aspect = window_size_.x / window_size_.y;
    persp_matrix_ = Mat4::perspective_rh(vfov_ * 3.14159265358979323846f / 180.0f, aspect, z_near_, z_far_);
}
void Camera::look_at(Vec3 eye, Vec3 center, Vec3 up)
{
    position_ = eye;
    center_ = center;
    up_ = up;
    update_view();
}
void Camera::set_window_size(Vec2 w)
{
    window_size_ = w;
    update_persp();
}

// ---- load_scene (src/scene/mod.rs:106-295) -----------------------------------
namespace {

bool find_mesh(const gltf::Document &d, int64_t node, std::vector<Mat4> &transforms, int64_t mesh_index)
{
    const gltf::Node &n = d.nodes[(size_t)node];
    transforms.push_back(Mat4::from_cols_array(n.matrix));
    if (n.mesh == mesh_index) return true;
    for (int64_t c : n.children)
        if (c >= 0 && (size_t)c < d.nodes.size() && find_mesh(d, c, transforms, mesh_index)) return true;
    transforms.pop_back();
    return false;
}

Mat4 calc_mesh_global_transform(const gltf::Document &d, int64_t mesh_index)
{
    Mat4 global = Mat4::identity();
    std::vector<Mat4> transforms;
    for (size_t i = 0; i < d.nodes.size(); ++i) {
        if (find_mesh(d, (int64_t)i, transforms, mesh_index)) {
            for (const Mat4 &t : transforms) global = global * t;
            break;
        }
    }
    return global;
}

}  // namespace

HostScene decode_scene(const std::string &path)
{
    gltf::Document d;
    try {
        d = gltf::load(path);
    } catch (const std::exception &e) {
        throw Error(SOLHIP_ERROR_INVALID_ARGUMENT, std::string("load_scene: ") + e.what());
    }
    HostScene hs;
    for (const gltf::Material &m : d.materials) {
        MaterialInfo mi;
        mi.base_color = {m.base_color[0], m.base_color[1], m.base_color[2], m.base_color[3]};
        mi.metallic_factor = m.metallic;
        mi.roughness_factor = m.roughness;
        mi.emissive_factor = {m.emissive[0], m.emissive[1], m.emissive[2]};
        hs.materials.push_back(mi);
    }
    for (size_t mi = 0; mi < d.meshes.size(); ++mi) {
        HostMesh mesh;
        mesh.name = d.mesh_names[mi];
        std::vector<ModelVertex> &verts = mesh.vertices;
        std::vector<uint32_t> &inds = mesh.indices;
        for (size_t pi = 0; pi < d.meshes[mi].size(); ++pi) {
            const gltf::Primitive &p = d.meshes[mi][pi];
            const size_t offset = verts.size();
            if (p.position >= 0) {
                const std::vector<float> pos = d.read_floats(p.position);
                const std::vector<float> nrm = d.read_floats(p.normal);
                const std::vector<float> uv = d.read_floats(p.texcoord0);
                const std::vector<float> col = d.read_floats(p.color0);
                const size_t cc = p.color0 >= 0 ? (size_t)d.accessors[(size_t)p.color0].ncomp : 4;
                const size_t n = pos.size() / 3;
                for (size_t k = 0; k < n; ++k) {
                    ModelVertex v;   // defaults: src/scene/mod.rs:182-193
                    v.pos = {pos[3 * k], pos[3 * k + 1], pos[3 * k + 2], 1.0f};
                    if (3 * k + 2 < nrm.size()) v.normal = {nrm[3 * k], nrm[3 * k + 1], nrm[3 * k + 2], 1.0f};
                    else v.normal = {0.0f, 1.0f, 0.0f, 1.0f};
                    if (2 * k + 1 < uv.size()) v.uv = {uv[2 * k], uv[2 * k + 1], 0.0f, 0.0f};
                    if (cc * k + cc - 1 < col.size())
                        v.color = {col[cc * k], col[cc * k + 1], col[cc * k + 2], cc == 4 ? col[cc * k + 3] : 1.0f};
                    verts.push_back(v);
                }
            }
            PrimitiveSection ps;
            ps.index = pi;
            ps.vertices = BufferPart{offset, verts.size() - offset};
            if (p.material >= 0) ps.material_index = (size_t)p.material;
            if (p.indices >= 0) {
                const size_t ioff = inds.size();
                const std::vector<uint32_t> ii = d.read_indices(p.indices);
                inds.insert(inds.end(), ii.begin(), ii.end());
                ps.indices = BufferPart{ioff, inds.size() - ioff};
            }
            mesh.primitive_sections.push_back(ps);
        }
        mesh.transform = calc_mesh_global_transform(d, (int64_t)mi);
        hs.meshes.push_back(std::move(mesh));
    }
    // first (default) camera only (src/scene/mod.rs:261-287)
    if (!d.cameras.empty() && d.cameras[0].perspective) {
        for (const gltf::Node &n : d.nodes) {
            if (n.camera == 0) {
                const gltf::Camera &c = d.cameras[0];
                hs.camera = Camera::from_view(Mat4::from_cols_array(n.matrix), c.yfov, c.znear,
                                              c.has_zfar ? c.zfar : 100.0f);
                break;
            }
        }
    }
    return hs;
}

Scene load_scene(std::shared_ptr<Context> ctx, const std::string &path)
{
    HostScene hs = decode_scene(path);
    Scene scene;
    scene.materials = hs.materials;
    scene.material_buffer = Buffer::from_data(ctx, scene.materials.empty() ? std::vector<MaterialInfo>(1)
                                                                           : scene.materials);
    for (HostMesh &hm : hs.meshes) {
        Mesh mesh;
        mesh.name = hm.name;
        mesh.transform = hm.transform;
        mesh.primitive_sections = hm.primitive_sections;
        if (!hm.indices.empty()) {
            mesh.index_buffer = Buffer::from_data(ctx, hm.indices);
            std::vector<uint64_t> wide(hm.indices.begin(), hm.indices.end());   // src/scene/mod.rs:228-233
            mesh.index_storage = Buffer::from_data(ctx, wide);
        }
        mesh.vertex_buffer = Buffer::from_data(ctx, hm.vertices.empty() ? std::vector<ModelVertex>(1) : hm.vertices);
        mesh.host_vertices = std::move(hm.vertices);
        mesh.host_indices = std::move(hm.indices);
        scene.meshes.push_back(std::move(mesh));
    }
    scene.camera = hs.camera;
    return scene;
}

}  // namespace scene

namespace ray {

// ---- BLAS / TLAS -------------------------------------------------------------
BLAS::BLAS(std::shared_ptr<Context> ctx, CommandBuffer &cmd, const std::vector<GeometryInstance> &geos, Mat4 transform,
           uint64_t vertex_stride, bool is_opaque)
    : ctx_(std::move(ctx)), transform_(transform)
{
    if (!is_opaque) throw Error(SOLHIP_ERROR_UNSUPPORTED, "BLAS::new: only opaque geometry is supported");
    std::vector<solhip_geometry> g;
    for (const GeometryInstance &gi : geos) {
        solhip_geometry s{};
        s.vertex_addr = gi.vertex_buffer;
        s.vertex_stride = vertex_stride;
        s.opaque = 1;
        s.first_vertex = gi.vertex_offset;
        s.vertex_count = gi.vertex_count;
        if (gi.index_buffer) {
            // indexed: BuildRangeInfo{primitive_count = index_count / 3, primitive_offset =
            // index byte offset, first_vertex = vertex_offset} (src/ray/acceleration.rs:182-200)
            s.index_addr = *gi.index_buffer;
            s.primitive_offset = *gi.index_offset_size;
            s.index_count = *gi.index_count;
        } else {
            // non-indexed: primitive_count = vertex_count / 3, primitive_offset = the
            // vertex byte offset, first_vertex = vertex_offset (acceleration.rs:185-188,
            // 194-199).  Vulkan applies both, so a primitive whose vertices start at
            // vertex_offset > 0 is read from 2 * vertex_offset on; kept as the reference has it.
            s.index_addr = 0;
            s.primitive_offset = gi.vertex_offset_size;
            s.index_count = 0;
        }
        g.push_back(s);
    }
    // build flags PREFER_FAST_TRACE (src/ray/acceleration.rs:213)
    check(solhip_blas_build(ctx_->handle(), cmd.stream, g.data(), (uint32_t)g.size(), SOLHIP_BUILD_PREFER_FAST_TRACE,
                            &blas_),
          "BLAS::new");
}
BLAS::~BLAS()
{
    if (blas_ && ctx_) solhip_blas_destroy(ctx_->handle(), blas_);
}
BLAS::BLAS(BLAS &&o) noexcept : ctx_(std::move(o.ctx_)), blas_(o.blas_), transform_(o.transform_) { o.blas_ = 0; }
BLAS &BLAS::operator=(BLAS &&o) noexcept
{
    if (this != &o) {
        if (blas_ && ctx_) solhip_blas_destroy(ctx_->handle(), blas_);
        ctx_ = std::move(o.ctx_);
        blas_ = o.blas_;
        transform_ = o.transform_;
        o.blas_ = 0;
    }
    return *this;
}

// TLAS::create_instances (src/ray/acceleration.rs:315-342): transposed
// transform (3x4 row-major), custom id i, mask 0xFF, sbt offset 0,
// FORCE_OPAQUE | TRIANGLE_FACING_CULL_DISABLE
std::vector<solhip_instance> TLAS::create_instances(const std::vector<BLAS> &blas)
{
    std::vector<solhip_instance> out(blas.size());
    for (size_t i = 0; i < blas.size(); ++i) {
        const Mat4 t = blas[i].get_transform().transpose();
        std::memcpy(out[i].transform, t.data(), 12 * sizeof(float));
        out[i].id_mask = ((uint32_t)i & 0xFFFFFFu) | (0xFFu << 24);
        out[i].offset_flags = (blas[i].hit_group_index() & 0xFFFFFFu) | ((0x4u | 0x1u) << 24);
        out[i].blas = blas[i].handle();
    }
    return out;
}

TLAS::TLAS(std::shared_ptr<Context> ctx, CommandBuffer &cmd, const std::vector<BLAS> &blas) : ctx_(std::move(ctx))
{
    const std::vector<solhip_instance> inst = create_instances(blas);
    check(solhip_tlas_build(ctx_->handle(), cmd.stream, inst.data(), (uint32_t)inst.size(), &tlas_), "TLAS::new");
}
TLAS::~TLAS()
{
    if (tlas_ && ctx_) solhip_tlas_destroy(ctx_->handle(), tlas_);
}
TLAS::TLAS(TLAS &&o) noexcept : ctx_(std::move(o.ctx_)), tlas_(o.tlas_) { o.tlas_ = 0; }
TLAS &TLAS::operator=(TLAS &&o) noexcept
{
    if (this != &o) {
        if (tlas_ && ctx_) solhip_tlas_destroy(ctx_->handle(), tlas_);
        ctx_ = std::move(o.ctx_);
        tlas_ = o.tlas_;
        o.tlas_ = 0;
    }
    return *this;
}
void TLAS::regenerate(CommandBuffer &cmd, const std::vector<BLAS> &blas)
{
    const std::vector<solhip_instance> inst = create_instances(blas);
    check(solhip_tlas_rebuild(ctx_->handle(), cmd.stream, tlas_, inst.data(), (uint32_t)inst.size()),
          "TLAS::regenerate");
}

// ---- SceneDescription (src/ray/mod.rs:49-206) -------------------------------
SceneDescription SceneDescription::from_scene(std::shared_ptr<Context> ctx, const scene::Scene &scene)
{
    std::vector<const scene::Mesh *> meshes;
    std::vector<Mat4> transforms;
    for (const scene::Mesh &m : scene.meshes) {
        meshes.push_back(&m);
        transforms.push_back(m.transform);
    }
    return from_meshes(std::move(ctx), meshes, transforms, &scene.material_buffer);
}

SceneDescription SceneDescription::from_meshes(std::shared_ptr<Context> ctx,
                                               const std::vector<const scene::Mesh *> &meshes,
                                               const std::vector<Mat4> &mesh_transforms, const Buffer *material_buffer)
{
    SceneDescription sd;
    CommandBuffer cmd;
    for (size_t i = 0; i < meshes.size(); ++i) {
        const scene::Mesh &mesh = *meshes[i];
        for (const scene::PrimitiveSection &prim : mesh.primitive_sections) {
            if (!prim.indices) continue;   // non-indexed primitives are not supported
            GeometryInstance gi;
            gi.vertex_buffer = mesh.vertex_buffer.get_device_address();
            gi.vertex_count = prim.get_vertex_count();
            gi.vertex_offset_size = prim.get_vertex_offset_size();
            gi.vertex_offset = prim.get_vertex_offset();
            if (mesh.index_buffer) {
                gi.index_buffer = mesh.index_buffer->get_device_address();
                gi.index_count = prim.get_index_count();
                gi.index_offset_size = prim.get_index_offset_size<uint32_t>();
            }
            sd.vertex_descriptors_.push_back(prim.get_vertex_descriptor(mesh.vertex_buffer));
            if (mesh.index_storage) sd.index_descriptors_.push_back(prim.get_index_descriptor<uint64_t>(*mesh.index_storage));
            if (material_buffer && prim.material_index)
                sd.mat_descriptors_.push_back(prim.get_material_descriptor(*material_buffer));
            else if (material_buffer)
                sd.mat_descriptors_.push_back({material_buffer->get_device_address(), 0, sizeof(scene::MaterialInfo)});
            SceneInstance inst;
            inst.id = (uint32_t)sd.instances_.size();
            inst.transform = mesh_transforms[i];
            inst.transform_it = mesh_transforms[i].inverse().transpose();
            sd.instances_.push_back(inst);
            sd.blas_.emplace_back(ctx, cmd, std::vector<GeometryInstance>{gi}, mesh_transforms[i],
                                  scene::ModelVertex::stride(), true);
            sd.blas_to_instances_.emplace_back(i, std::vector<size_t>{inst.id});
        }
    }
    sd.tlas_ = std::make_unique<TLAS>(ctx, cmd, sd.blas_);
    sd.instances_buffer_ = Buffer::from_data(ctx, sd.instances_.empty() ? std::vector<SceneInstance>(1) : sd.instances_);
    sd.build_tables(ctx);
    return sd;
}

void SceneDescription::build_tables(std::shared_ptr<Context> ctx)
{
    std::vector<uint64_t> v, ix, m;
    for (const auto &d : vertex_descriptors_) v.push_back(d.address + d.offset);
    for (const auto &d : index_descriptors_) ix.push_back(d.address + d.offset);
    for (const auto &d : mat_descriptors_) m.push_back(d.address + d.offset);
    if (v.empty()) v.push_back(0);
    if (ix.empty()) ix.push_back(0);
    vertex_table_ = Buffer::from_data(ctx, v);
    index_table_ = Buffer::from_data(ctx, ix);
    tables_ = solhip_scene_tables{};
    tables_.instances_addr = instances_buffer_.get_device_address();
    tables_.instance_count = (uint32_t)instances_.size();
    tables_.index_size = 8;
    tables_.vertex_table_addr = vertex_table_.get_device_address();
    tables_.index_table_addr = index_table_.get_device_address();
    if (!m.empty()) {
        material_table_ = Buffer::from_data(ctx, m);
        tables_.material_table_addr = material_table_.get_device_address();
    }
    tables_.object_count = (uint32_t)vertex_descriptors_.size();
}

// src/ray/mod.rs:162-176; the map is keyed like the reference (mesh index)
void SceneDescription::blas_transform(const Mat4 &transform, size_t index)
{
    blas_.at(index).set_transform(transform);
    for (const auto &kv : blas_to_instances_)
        if (kv.first == index)
            for (size_t ii : kv.second) instances_.at(ii).update_transform(transform);
}
void SceneDescription::blas_transforms(const std::vector<Mat4> &transforms)
{
    for (size_t i = 0; i < transforms.size(); ++i) blas_transform(transforms[i], i);
}
void SceneDescription::tlas_regenerate(CommandBuffer &cmd) { tlas_->regenerate(cmd, blas_); }
void SceneDescription::update() { instances_buffer_.update(instances_); }

// ---- Pipeline / SBT -----------------------------------------------------------
namespace {
bool ends_with(const std::string &s, const std::string &suffix)
{
    return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}
}  // namespace

Pipeline::Pipeline(std::shared_ptr<Context> ctx, const PipelineInfo &info, const solhip_spec *override_spec)
    : ctx_(std::move(ctx))
{
    // the raygen shader names the program (assets/glsl/{ao,pathtrace,debug}.rgen;
    // "pathtrace_wavefront.rgen" selects the wavefront form of pathtrace.*,
    // whose extensions come in override_spec)
    for (const auto &s : info.shaders) {
        if (s.second != ShaderStage::Raygen) continue;
        if (ends_with(s.first, "pathtrace_wavefront.rgen")) kind_ = SOLHIP_PROG_PATHTRACE;
        else if (ends_with(s.first, "pathtrace.rgen")) kind_ = SOLHIP_PROG_PATHTRACE_GLSL;
        else if (ends_with(s.first, "ao.rgen")) kind_ = SOLHIP_PROG_AO_GLSL;
        else if (ends_with(s.first, "debug.rgen")) kind_ = SOLHIP_PROG_DEBUG;
        else if (ends_with(s.first, "ao_plain.rgen")) kind_ = SOLHIP_PROG_AO;
    }
    if (!kind_) throw Error(SOLHIP_ERROR_INVALID_ARGUMENT, "ray::Pipeline::new: unknown raygen shader set");
    solhip_spec_default(kind_, &spec_);
    if (override_spec) spec_ = *override_spec;
    if (!info.specialization_data.empty()) spec_.enable_sky = info.specialization_data[0];
    check(solhip_program_create(ctx_->handle(), kind_, &spec_, &prog_), "ray::Pipeline::new");
}
Pipeline::~Pipeline()
{
    if (prog_ && ctx_) solhip_program_destroy(ctx_->handle(), prog_);
}

ShaderBindingTable::ShaderBindingTable(std::shared_ptr<Context> ctx, const Pipeline &pipeline,
                                       const ShaderBindingTableInfo &info)
    : ctx_(std::move(ctx)), pipeline_(&pipeline)
{
    if (info.raygen_indices.empty()) throw Error(SOLHIP_ERROR_INVALID_ARGUMENT, "ShaderBindingTable: no raygen group");
}

void ShaderBindingTable::cmd_trace_rays(CommandBuffer &cmd, Extent3D extent) const
{
    const Pipeline *p = cmd.pipeline ? cmd.pipeline : pipeline_;
    if (!cmd.scene) throw Error(SOLHIP_ERROR_INVALID_ARGUMENT, "cmd_trace_rays: no scene bound");
    if (!cmd.images.accum) throw Error(SOLHIP_ERROR_INVALID_ARGUMENT, "cmd_trace_rays: no target image bound");
    solhip_images im{};
    im.accum = cmd.images.accum->handle();
    im.out_rgba8 = cmd.images.render ? cmd.images.render->handle() : 0;
    im.blue_noise = cmd.images.blue_noise ? cmd.images.blue_noise->handle() : 0;
    im.blue_w = cmd.images.blue_w;
    im.blue_h = cmd.images.blue_h;
    check(solhip_trace(ctx_->handle(), cmd.stream, p->handle(), cmd.scene->tlas().handle(), &cmd.scene->tables(),
                       &cmd.uniforms, cmd.accum_start_frame, &im, extent.width, extent.height),
          "cmd_trace_rays");
}

}  // namespace ray
}  // namespace sol
