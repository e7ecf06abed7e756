"""Byte layouts shared with the sol-rs Rust API (numpy dtypes).

Each dtype reproduces a `#[repr(C)]` record of the reference so that buffers
built for the Vulkan path can be handed to the HIP path unchanged.
"""
import numpy as np

# src/scene/mesh.rs:7-25  ModelVertex {pos, color, normal, uv}: 4 x vec4 = 64 B
MODEL_VERTEX = np.dtype([("pos", "<f4", 4), ("color", "<f4", 4), ("normal", "<f4", 4), ("uv", "<f4", 4)])

# src/scene/mod.rs:19-29  MaterialInfo: 48 B
MATERIAL = np.dtype([("base_color", "<f4", 4), ("emissive", "<f4", 3), ("padding0", "<f4"),
                     ("metallic", "<f4"), ("roughness", "<f4"), ("padding1", "<f4"), ("padding2", "<f4")])

# src/ray/mod.rs:16-24  SceneInstance: 144 B (column-major mat4, as glam stores them)
SCENE_INSTANCE = np.dtype([("id", "<u4"), ("texture_offset", "<u4"), ("padding", "<f4", 2),
                           ("transform", "<f4", 16), ("transform_it", "<f4", 16)])

# src/ray/acceleration.rs:256-263  InstanceDescriptor (VkAccelerationStructureInstanceKHR): 64 B.
# transform = row-major 3x4 object-to-world; id_mask = id:24 | mask:8 << 24;
# offset_flags = sbt offset:24 | flags:8 << 24; blas = BLAS handle (the
# reference stores the BLAS device address here).
INSTANCE_DESC = np.dtype([("transform", "<f4", 12), ("id_mask", "<u4"), ("offset_flags", "<u4"),
                          ("blas", "<u8")])

# examples/{4-ray-ao,5-pathtrace}.rs:9-17  SceneUniforms: 6 x mat4 + uvec3 (400 B)
UNIFORMS = np.dtype([("model", "<f4", 16), ("view", "<f4", 16), ("view_inverse", "<f4", 16),
                     ("projection", "<f4", 16), ("projection_inverse", "<f4", 16),
                     ("model_view_projection", "<f4", 16), ("frame", "<u4", 3), ("pad", "<u4")])

# raw ray / hit records of the trace_rays entry point
RAY = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
# hit: geom = gl_GeometryIndexEXT, prim = gl_PrimitiveID (within the geometry)
HIT = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("inst", "<u4"), ("prim", "<u4"), ("geom", "<u4")])
MISS = 0xFFFFFFFF

# VkGeometryInstanceFlagBitsKHR used by TLAS::create_instances
# (src/ray/acceleration.rs:337)
GEOMETRY_INSTANCE_TRIANGLE_FACING_CULL_DISABLE = 0x1
GEOMETRY_INSTANCE_FORCE_OPAQUE = 0x4

assert MODEL_VERTEX.itemsize == 64
assert MATERIAL.itemsize == 48
assert SCENE_INSTANCE.itemsize == 144
assert INSTANCE_DESC.itemsize == 64
assert UNIFORMS.itemsize == 400
assert RAY.itemsize == 32
assert HIT.itemsize == 24
