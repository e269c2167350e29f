//! Rust binding of `include/tinytrace.h`, the C ABI of the MI355X path.
//!
//! NOT COMPILED HERE: the build image has no Rust toolchain.  The struct and
//! function declarations are checked against the C header by
//! `tests/test_rust_binding_cpu.py` (field order, types, and the offsets the
//! `#[repr(C)]` rules give against gcc's `offsetof`).
//!
//! Where the reference calls its Vulkan backend this crate's functions take
//! over:
//!
//! * `Context::new` (crates/tinytrace_backend/src/lib.rs:171-213) ->
//!   [`tt_context_create`];
//! * `scene::Scene::new` + the driver's BLAS/TLAS build (src/scene.rs:28-186)
//!   -> [`tt_scene_create`];
//! * `Integrator::integrate` (src/integrate/mod.rs:39-61), dispatched from
//!   `Renderer::render_to_target` (src/lib.rs:150-155) -> [`tt_render`];
//! * `Context::download` (crates/tinytrace_backend/src/copy.rs:183-298) ->
//!   [`tt_target_download`].
//!
//! The byte layouts marked "reference layout" are those of the reference's own
//! `#[repr(C)]` types (`Constants`, `ProcessedVertex`, `ProcessedMesh`,
//! `MaterialConstants`, `ProcessedMaterial`, `Instance`, `EmissiveTriangle`),
//! so the reference can pass its processed arrays with `bytemuck::cast_slice`.
//! The safe layer at the end (`Context`, `Scene`, `Target`) is what
//! `Renderer` would hold instead of its backend handles.
#![allow(non_camel_case_types)]

use std::ffi::CStr;
use std::os::raw::{c_char, c_int};

// ------------------------------------------------------------------ status --

pub type tt_status = c_int;
pub const TT_OK: tt_status = 0;
pub const TT_NO_DEVICE: tt_status = 1;
pub const TT_HIP_ERROR: tt_status = 2;
pub const TT_KERNEL_LOAD: tt_status = 3;
pub const TT_INVALID_ARGUMENT: tt_status = 4;
pub const TT_STALE_HANDLE: tt_status = 5;
pub const TT_OUT_OF_MEMORY: tt_status = 6;
pub const TT_OVER_BUDGET: tt_status = 7;
pub const TT_UNSUPPORTED: tt_status = 8;
pub const TT_INTERNAL_ERROR: tt_status = 9;

pub const TT_LIGHT_SAMPLING_ON_HIT: u32 = 1;
pub const TT_LIGHT_SAMPLING_NEXT_EVENT_ESTIMATION: u32 = 2;
pub const TT_LIGHT_SAMPLING_RESTIR: u32 = 3;
pub const TT_INVALID_INDEX: u32 = 0xffff_ffff;
pub const TT_BVH_BUILDER_HOST_SAH: u32 = 0;
pub const TT_BVH_BUILDER_DEVICE_LBVH: u32 = 1;
pub const TT_RNG_REFERENCE: u32 = 0;
pub const TT_RNG_PCG: u32 = 1;
pub const TT_RENDER_COLLECT_STATS: u32 = 1;
pub const TT_REQUEST_QUEUED: u32 = 0;
pub const TT_REQUEST_ADMITTED: u32 = 1;
pub const TT_REQUEST_DONE: u32 = 2;
pub const TT_TRACE_CULL_BACK_FACES: u32 = 1;
pub const TT_TRACE_STATIC_MAX_RAYS: u32 = 65536;
pub const TT_TEXTURE_FORMAT_RGBA8_SRGB: u32 = 0;
pub const TT_TEXTURE_FORMAT_RG8_UNORM: u32 = 1;
pub const TT_TEXTURE_FORMAT_R8_UNORM: u32 = 2;
pub const TT_TEXTURE_FORMAT_R16_FLOAT: u32 = 3;
pub const TT_TEXTURE_FORMAT_BC1_SRGB: u32 = 4;
pub const TT_TEXTURE_FORMAT_BC5_UNORM: u32 = 5;
// Texture channels: the fields of MaterialTextures in order
// (src/asset/material.rs:203-231), i.e. the order of tt_material::textures.
pub const TT_TEXTURE_BASE_COLOR: u32 = 0;
pub const TT_TEXTURE_SPECULAR_COLOR: u32 = 1;
pub const TT_TEXTURE_TRANSMISSION_COLOR: u32 = 2;
pub const TT_TEXTURE_COAT_COLOR: u32 = 3;
pub const TT_TEXTURE_FUZZ_COLOR: u32 = 4;
pub const TT_TEXTURE_EMISSION_COLOR: u32 = 5;
pub const TT_TEXTURE_GEOMETRY_NORMAL: u32 = 6;
pub const TT_TEXTURE_BASE_WEIGHT: u32 = 7;
pub const TT_TEXTURE_BASE_METALNESS: u32 = 8;
pub const TT_TEXTURE_BASE_DIFFUSE_ROUGHNESS: u32 = 9;
pub const TT_TEXTURE_SPECULAR_WEIGHT: u32 = 10;
pub const TT_TEXTURE_SPECULAR_ROUGHNESS: u32 = 11;
pub const TT_TEXTURE_SPECULAR_ROUGHNESS_ANISOTROPY: u32 = 12;
pub const TT_TEXTURE_SPECULAR_IOR: u32 = 13;
pub const TT_TEXTURE_SPECULAR_ROTATION: u32 = 14;
pub const TT_TEXTURE_TRANSMISSION_WEIGHT: u32 = 15;
pub const TT_TEXTURE_TRANSMISSION_DEPTH: u32 = 16;
pub const TT_TEXTURE_COAT_WEIGHT: u32 = 17;
pub const TT_TEXTURE_COAT_ROUGHNESS: u32 = 18;
pub const TT_TEXTURE_COAT_ROUGHNESS_ANISOTROPY: u32 = 19;
pub const TT_TEXTURE_COAT_IOR: u32 = 20;
pub const TT_TEXTURE_COAT_DARKENING: u32 = 21;
pub const TT_TEXTURE_COAT_ROTATION: u32 = 22;
pub const TT_TEXTURE_FUZZ_WEIGHT: u32 = 23;
pub const TT_TEXTURE_FUZZ_ROUGHNESS: u32 = 24;
pub const TT_TEXTURE_EMISSION_LUMINANCE: u32 = 25;
pub const TT_TEXTURE_GEOMETRY_OPACITY: u32 = 26;
pub const TT_TEXTURE_CHANNELS: u32 = 27;

// -------------------------------------------------- reference byte layouts --

/// `Constants` (src/lib.rs:234-250), 368 B, column-major matrices.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_constants {
    pub view: [f32; 16],
    pub proj: [f32; 16],
    pub proj_view: [f32; 16],
    pub inverse_view: [f32; 16],
    pub inverse_proj: [f32; 16],
    pub camera_position: [f32; 4],
    pub frame_index: u32,
    pub accumulated_frame_count: u32,
    pub sample_count: u32,
    pub bounce_count: u32,
    pub screen_size: [u32; 2],
    pub tonemap: u32,
    pub light_sampling: u32,
}

/// `ProcessedVertex` (src/asset/mod.rs:42-47), 8 B.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_vertex {
    pub texture_coordinates: [u16; 2],
    pub tangent_frame: u32,
}

/// `ProcessedMesh` (src/asset/mod.rs:49-64), 32 B.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_mesh_record {
    pub center: [f32; 3],
    pub radius: f32,
    pub vertex_offset: u32,
    pub vertex_count: u32,
    pub index_offset: u32,
    pub index_count: u32,
}

/// `MaterialConstants` (src/asset/material.rs:203-231), 192 B.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_material_constants {
    pub base_color: [f32; 4],
    pub specular_color: [f32; 4],
    pub transmission_color: [f32; 4],
    pub coat_color: [f32; 4],
    pub fuzz_color: [f32; 4],
    pub emission_color: [f32; 4],
    pub geometry_normal: [f32; 4],
    pub base_weight: f32,
    pub base_metalness: f32,
    pub base_diffuse_roughness: f32,
    pub specular_weight: f32,
    pub specular_roughness: f32,
    pub specular_roughness_anisotropy: f32,
    pub specular_ior: f32,
    pub specular_rotation: f32,
    pub transmission_weight: f32,
    pub transmission_depth: f32,
    pub coat_weight: f32,
    pub coat_roughness: f32,
    pub coat_roughness_anisotropy: f32,
    pub coat_ior: f32,
    pub coat_darkening: f32,
    pub coat_rotation: f32,
    pub fuzz_weight: f32,
    pub fuzz_roughness: f32,
    pub emission_luminance: f32,
    pub geometry_opacity: f32,
}

/// `ProcessedMaterial` (src/asset/material.rs:268-274), 304 B.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct tt_material {
    pub constants: tt_material_constants,
    pub textures: [u32; 27],
    pub padding: u32,
}

/// `Instance` (src/scene.rs:236-246), 208 B.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct tt_instance {
    pub transform: [f32; 16],
    pub inverse_transform: [f32; 16],
    pub normal_transform: [f32; 16],
    pub vertex_offset: u32,
    pub index_offset: u32,
    pub material_index: u32,
    pub mesh_index: u32,
}

/// `EmissiveTriangle` (src/asset/mod.rs:66-73), 36 B.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_emissive_triangle {
    pub positions: [[i16; 3]; 3],
    pub hash: u16,
    pub texture_coordinates: [[u16; 2]; 3],
    pub instance_index: u32,
}

/// Everything `scene::Scene::new` uploads (src/scene.rs:28-186), flattened.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct tt_scene_desc {
    pub positions: *const i16,
    pub position_count: u32,
    pub vertices: *const tt_vertex,
    pub vertex_count: u32,
    pub indices: *const u32,
    pub index_count: u32,
    pub meshes: *const tt_mesh_record,
    pub mesh_count: u32,
    pub materials: *const tt_material,
    pub material_count: u32,
    pub instances: *const tt_instance,
    pub instance_count: u32,
    pub emissive_triangles: *const tt_emissive_triangle,
    pub emissive_triangle_count: u32,
    pub textures: *const tt_texture_record,
    pub texture_count: u32,
    pub texture_data: *const u8,
    pub texture_data_bytes: u64,
}

/// One processed texture's mip chain in `tt_scene_desc::texture_data`.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_texture_record {
    pub format: u32,
    pub width: u32,
    pub height: u32,
    pub levels: u32,
    pub offsets: [u32; 16],
}

// ----------------------------------------------------------------- handles --

#[repr(C)]
pub struct tt_context {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct tt_scene {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct tt_target {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct tt_server {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct tt_mlp {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct tt_asset_scene {
    _opaque: [u8; 0],
}

// ------------------------------------------------------- library's own types --

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_context_desc {
    pub device: i32,
    pub flags: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_scene_options {
    pub bvh_builder: u32,
    pub reserved: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_scene_info {
    pub triangle_count: u32,
    pub unique_triangle_count: u32,
    pub node_count: u32,
    pub max_depth: u32,
    pub sah_cost: f64,
    pub expected_bytes_per_ray: f64,
    pub build_ms: f64,
    pub device_bytes: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_render_settings {
    pub rng: u32,
    pub seed: u32,
    pub restir_candidates: u32,
    pub restir_temporal: u32,
    pub restir_spatial: u32,
    pub flags: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_render_stats {
    pub paths: u64,
    pub extension_rays: u64,
    pub shadow_rays: u64,
    pub extend_launches: u32,
    pub shadow_launches: u32,
    pub extend_ms: f64,
    pub shadow_ms: f64,
    pub integrate_ms: f64,
    pub restir_rays: u64,
    pub restir_ms: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_server_desc {
    pub hbm_budget_bytes: u64,
    pub pool_paths: u32,
    pub batches_in_flight: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_server_request {
    pub scene: u32,
    pub rng: u32,
    pub seed: u32,
    pub passes: u32,
    pub constants: tt_constants,
    pub restir_candidates: u32,
    pub restir_temporal: u32,
    pub restir_spatial: u32,
    pub reserved: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_request_status {
    pub state: u32,
    pub passes_done: u32,
    pub submitted_ns: u64,
    pub admitted_ns: u64,
    pub done_ns: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_server_stats {
    pub batches: u64,
    pub paths: u64,
    pub queued: u64,
    pub active: u64,
    pub done: u64,
    pub budget_used_bytes: u64,
    pub busy_ms: f64,
    pub cache_queries: u64,
    pub cache_samples: u64,
    pub judge_samples: u64,
    pub judge_traced: f64,
    pub judge_cached: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_mlp_desc {
    pub seed: u64,
    pub learning_rate: f32,
    pub reserved: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_cache_settings {
    pub train_samples: u32,
    pub train_steps: u32,
    pub judge_fraction: f32,
    pub records_per_path: u32,
    pub ordered_training: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_cache_stats {
    pub queries: u32,
    pub samples: u32,
    pub loss: f64,
    pub judge_samples: u32,
    pub reserved: u32,
    pub judge_traced: f64,
    pub judge_cached: f64,
    pub judge_loss: f64,
    pub infer_ms: f64,
    pub train_ms: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_ray {
    pub origin: [f32; 3],
    pub tmin: f32,
    pub direction: [f32; 3],
    pub tmax: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_hit {
    pub t: f32,
    pub instance: u32,
    pub primitive: u32,
    pub barycentrics: [f32; 2],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_bsdf_query {
    pub wi: [f32; 3],
    pub wo: [f32; 3],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_bsdf_result {
    pub bsdf: [f32; 3],
    pub density: f32,
}

/// `Camera` (src/camera.rs:6-26).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_camera {
    pub position: [f32; 3],
    pub yaw: f32,
    pub pitch: f32,
    pub z_near: f32,
    pub z_far: f32,
    pub fov: f32,
}

/// `Config` (src/config.rs:23-43).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_config {
    pub bounce_count: u32,
    pub sample_count: u32,
    pub tonemap: u32,
    pub light_sampling: u32,
}

/// `asset::Mesh` (src/asset/mod.rs:20-27).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct tt_mesh_desc {
    pub positions: *const f32,
    pub texture_coordinates: *const f32,
    pub indices: *const u32,
    pub normals: *const f32,
    pub tangents: *const f32,
    pub vertex_count: u32,
    pub index_count: u32,
}

/// `asset::Material` with constant parameters (src/asset/material.rs:8-175).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct tt_material_desc {
    pub base_weight: f32,
    pub base_color: [f32; 3],
    pub base_metalness: f32,
    pub base_diffuse_roughness: f32,
    pub specular_weight: f32,
    pub specular_color: [f32; 3],
    pub specular_roughness: f32,
    pub specular_roughness_anisotropy: f32,
    pub specular_ior: f32,
    pub specular_rotation: f32,
    pub transmission_weight: f32,
    pub transmission_color: [f32; 3],
    pub transmission_depth: f32,
    pub coat_weight: f32,
    pub coat_color: [f32; 3],
    pub coat_roughness: f32,
    pub coat_roughness_anisotropy: f32,
    pub coat_ior: f32,
    pub coat_darkening: f32,
    pub coat_rotation: f32,
    pub fuzz_weight: f32,
    pub fuzz_color: [f32; 3],
    pub fuzz_roughness: f32,
    pub emission_luminance: f32,
    pub emission_color: [f32; 3],
    pub geometry_opacity: f32,
    pub geometry_normal: [f32; 3],
}

/// A material parameter given as a texture (`Param::Texture`).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct tt_texture_desc {
    pub channel: u32,
    pub width: u32,
    pub height: u32,
    pub texels: *const f32,
}

// --------------------------------------------------------------- functions --

#[link(name = "tinytrace_amd")]
extern "C" {
    pub fn tt_context_create(desc: *const tt_context_desc, out: *mut *mut tt_context) -> tt_status;
    pub fn tt_context_destroy(context: *mut tt_context);
    pub fn tt_last_error(context: *const tt_context) -> *const c_char;

    pub fn tt_scene_create(context: *mut tt_context, desc: *const tt_scene_desc, out: *mut *mut tt_scene) -> tt_status;
    pub fn tt_scene_create_ex(context: *mut tt_context, desc: *const tt_scene_desc, options: *const tt_scene_options, out: *mut *mut tt_scene) -> tt_status;
    pub fn tt_scene_destroy(scene: *mut tt_scene);
    pub fn tt_scene_info_get(scene: *const tt_scene, out: *mut tt_scene_info) -> tt_status;

    pub fn tt_target_create(context: *mut tt_context, width: u32, height: u32, out: *mut *mut tt_target) -> tt_status;
    pub fn tt_target_destroy(target: *mut tt_target);
    pub fn tt_render(context: *mut tt_context, scene: *const tt_scene, constants: *const tt_constants, target: *mut tt_target) -> tt_status;
    pub fn tt_render_ex(context: *mut tt_context, scene: *const tt_scene, constants: *const tt_constants, settings: *const tt_render_settings, target: *mut tt_target) -> tt_status;
    pub fn tt_target_download(context: *mut tt_context, target: *const tt_target, rgba: *mut f32, bytes: usize) -> tt_status;
    pub fn tt_target_display(context: *mut tt_context, target: *const tt_target, constants: *const tt_constants, rgba8: *mut u8, bytes: usize) -> tt_status;
    pub fn tt_timing_ms(context: *mut tt_context, name: *const c_char, out: *mut f64) -> tt_status;
    pub fn tt_render_stats_get(context: *mut tt_context, out: *mut tt_render_stats) -> tt_status;

    pub fn tt_server_create(context: *mut tt_context, desc: *const tt_server_desc, out: *mut *mut tt_server) -> tt_status;
    pub fn tt_server_destroy(server: *mut tt_server);
    pub fn tt_server_add_scene(server: *mut tt_server, scene: *const tt_scene, slot: *mut u32) -> tt_status;
    pub fn tt_server_submit(server: *mut tt_server, request: *const tt_server_request, id: *mut u64) -> tt_status;
    pub fn tt_server_step(server: *mut tt_server, completed: *mut u32) -> tt_status;
    pub fn tt_server_status(server: *mut tt_server, id: u64, out: *mut tt_request_status) -> tt_status;
    pub fn tt_server_take_result(server: *mut tt_server, id: u64, rgba: *mut f32, bytes: usize) -> tt_status;
    pub fn tt_server_stats_get(server: *mut tt_server, out: *mut tt_server_stats) -> tt_status;

    pub fn tt_mlp_create(context: *mut tt_context, desc: *const tt_mlp_desc, out: *mut *mut tt_mlp) -> tt_status;
    pub fn tt_mlp_destroy(mlp: *mut tt_mlp);
    pub fn tt_mlp_set_bounds(mlp: *mut tt_mlp, lo: *const f32, hi: *const f32) -> tt_status;
    pub fn tt_mlp_params(mlp: *mut tt_mlp, get: *mut f32, set: *const f32, count: u32) -> tt_status;
    pub fn tt_mlp_run(mlp: *mut tt_mlp, op: c_int, records: *const f32, targets: *const f32, n: u32, out: *mut f32, loss: *mut f64) -> tt_status;
    pub fn tt_render_cached(context: *mut tt_context, scene: *const tt_scene, constants: *const tt_constants, settings: *const tt_render_settings, cache: *mut tt_mlp, cache_settings: *const tt_cache_settings, target: *mut tt_target, stats: *mut tt_cache_stats) -> tt_status;
    pub fn tt_server_set_cache(server: *mut tt_server, scene_slot: u32, cache: *mut tt_mlp, settings: *const tt_cache_settings) -> tt_status;

    pub fn tt_trace_closest(context: *mut tt_context, scene: *const tt_scene, rays: *const tt_ray, hits: *mut tt_hit, count: u32, flags: u32) -> tt_status;
    pub fn tt_trace_closest_device(context: *mut tt_context, scene: *const tt_scene, device_rays: *const tt_ray, device_hits: *mut tt_hit, count: u32, flags: u32) -> tt_status;
    pub fn tt_trace_closest_static(context: *mut tt_context, scene: *const tt_scene, rays: *const tt_ray, hits: *mut tt_hit, count: u32, flags: u32) -> tt_status;
    pub fn tt_bsdf_evaluate(context: *mut tt_context, material: *const tt_material_constants, queries: *const tt_bsdf_query, results: *mut tt_bsdf_result, count: u32, rng_state: u32) -> tt_status;
    pub fn tt_debug_math(context: *mut tt_context, op: c_int, x: *const f32, y: *const f32, out: *mut f32, count: u32) -> tt_status;

    pub fn tt_camera_default(out: *mut tt_camera);
    pub fn tt_config_default(out: *mut tt_config);
    pub fn tt_constants_build(camera: *const tt_camera, config: *const tt_config, width: u32, height: u32, frame_index: u32, accumulated_frame_count: u32, out: *mut tt_constants) -> tt_status;

    pub fn tt_asset_scene_create(out: *mut *mut tt_asset_scene) -> tt_status;
    pub fn tt_asset_scene_destroy(scene: *mut tt_asset_scene);
    pub fn tt_asset_set_block_compression(scene: *mut tt_asset_scene, enable: u32) -> tt_status;
    pub fn tt_asset_insert_mesh(scene: *mut tt_asset_scene, mesh: *const tt_mesh_desc, index: *mut u32) -> tt_status;
    pub fn tt_material_desc_default(out: *mut tt_material_desc);
    pub fn tt_asset_insert_material(scene: *mut tt_asset_scene, material: *const tt_material_desc, index: *mut u32) -> tt_status;
    pub fn tt_asset_insert_material_textured(scene: *mut tt_asset_scene, material: *const tt_material_desc, textures: *const tt_texture_desc, texture_count: u32, index: *mut u32) -> tt_status;
    pub fn tt_asset_add_node(scene: *mut tt_asset_scene, parent: u32, transform: *const f32, index: *mut u32) -> tt_status;
    pub fn tt_asset_add_model(scene: *mut tt_asset_scene, node: u32, mesh: u32, material: u32) -> tt_status;
    pub fn tt_asset_import_gltf(scene: *mut tt_asset_scene, path: *const c_char) -> tt_status;
    pub fn tt_asset_flatten(scene: *mut tt_asset_scene, out: *mut tt_scene_desc) -> tt_status;
}

// ------------------------------------------------------------- safe layer --

/// The library's failures, as the reference's `Error` enums name them
/// (src/error.rs:3-9, crates/tinytrace_backend/src/error.rs:7-28), with the
/// context's message.
#[derive(Debug, Clone, PartialEq, Eq)]
pub enum Error {
    NoSuitableDevice,
    Hip(String),
    KernelLoad(String),
    InvalidArgument(String),
    StaleHandle(String),
    OutOfMemory(String),
    OverBudget(String),
    Unsupported(String),
    Internal(String),
}

fn check(status: tt_status, context: *const tt_context) -> Result<(), Error> {
    if status == TT_OK {
        return Ok(());
    }
    let message = if context.is_null() {
        String::new()
    } else {
        // SAFETY: tt_last_error returns a NUL-terminated string owned by the
        // context, valid until its next call.
        unsafe { CStr::from_ptr(tt_last_error(context)) }.to_string_lossy().into_owned()
    };
    Err(match status {
        TT_NO_DEVICE => Error::NoSuitableDevice,
        TT_HIP_ERROR => Error::Hip(message),
        TT_KERNEL_LOAD => Error::KernelLoad(message),
        TT_INVALID_ARGUMENT => Error::InvalidArgument(message),
        TT_STALE_HANDLE => Error::StaleHandle(message),
        TT_OUT_OF_MEMORY => Error::OutOfMemory(message),
        TT_OVER_BUDGET => Error::OverBudget(message),
        TT_UNSUPPORTED => Error::Unsupported(message),
        _ => Error::Internal(message),
    })
}

/// Replaces the backend `Context` (crates/tinytrace_backend/src/lib.rs:171-213).
pub struct Context {
    raw: *mut tt_context,
}

impl Context {
    pub fn new(device: i32) -> Result<Self, Error> {
        let desc = tt_context_desc { device, flags: 0 };
        let mut raw = std::ptr::null_mut();
        check(unsafe { tt_context_create(&desc, &mut raw) }, std::ptr::null())?;
        Ok(Self { raw })
    }

    pub fn raw(&self) -> *mut tt_context {
        self.raw
    }

    /// `scene::Scene::new` (src/scene.rs:28-186): uploads the processed
    /// arrays and builds the BVH.
    pub fn create_scene(&mut self, desc: &tt_scene_desc) -> Result<Scene, Error> {
        let mut raw = std::ptr::null_mut();
        check(unsafe { tt_scene_create(self.raw, desc, &mut raw) }, self.raw)?;
        Ok(Scene { raw })
    }

    /// `create_render_target` (src/lib.rs:213-226).
    pub fn create_target(&mut self, width: u32, height: u32) -> Result<Target, Error> {
        let mut raw = std::ptr::null_mut();
        check(unsafe { tt_target_create(self.raw, width, height, &mut raw) }, self.raw)?;
        Ok(Target { raw, width, height })
    }

    /// `Integrator::integrate` (src/integrate/mod.rs:39-61) for one
    /// `render_to_target` call (src/lib.rs:124-161).
    pub fn integrate(&mut self, scene: &Scene, constants: &tt_constants, target: &mut Target) -> Result<(), Error> {
        check(unsafe { tt_render(self.raw, scene.raw, constants, target.raw) }, self.raw)
    }

    /// `render_to_texture`'s download (src/lib.rs:201-210): linear RGBA32F.
    pub fn download(&mut self, target: &Target) -> Result<Vec<[f32; 4]>, Error> {
        let mut pixels = vec![[0.0f32; 4]; target.width as usize * target.height as usize];
        let bytes = pixels.len() * std::mem::size_of::<[f32; 4]>();
        check(unsafe { tt_target_download(self.raw, target.raw, pixels.as_mut_ptr().cast(), bytes) }, self.raw)?;
        Ok(pixels)
    }

    /// `Renderer::timings` (src/lib.rs:167-178): "integrate", "post_process".
    pub fn timing_ms(&mut self, name: &CStr) -> Result<f64, Error> {
        let mut ms = 0.0;
        check(unsafe { tt_timing_ms(self.raw, name.as_ptr(), &mut ms) }, self.raw)?;
        Ok(ms)
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        // Scenes and targets may outlive the handle: the library frees the
        // context when its last child goes.
        unsafe { tt_context_destroy(self.raw) }
    }
}

pub struct Scene {
    raw: *mut tt_scene,
}

impl Drop for Scene {
    fn drop(&mut self) {
        unsafe { tt_scene_destroy(self.raw) }
    }
}

pub struct Target {
    raw: *mut tt_target,
    pub width: u32,
    pub height: u32,
}

impl Drop for Target {
    fn drop(&mut self) {
        unsafe { tt_target_destroy(self.raw) }
    }
}

// Size checks of the reference layouts, as the C header's static asserts.
const _: () = assert!(std::mem::size_of::<tt_constants>() == 368);
const _: () = assert!(std::mem::size_of::<tt_vertex>() == 8);
const _: () = assert!(std::mem::size_of::<tt_mesh_record>() == 32);
const _: () = assert!(std::mem::size_of::<tt_material_constants>() == 192);
const _: () = assert!(std::mem::size_of::<tt_material>() == 304);
const _: () = assert!(std::mem::size_of::<tt_instance>() == 208);
const _: () = assert!(std::mem::size_of::<tt_emissive_triangle>() == 36);
