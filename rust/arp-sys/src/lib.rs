//! Raw bindings of `libarp.so`, the MI355X-native particle step of ar-physics, declared in
//! `include/arp.h` (every function there has its declaration below, in the same order).
//!
//! The wire structs are byte-identical to the reference's std430 types
//! (`src/soft_body/particle_state/mod.rs:24-35`, `material_region/material.rs:11-28`,
//! `material_region/interaction.rs:10-15`, `material_region/region.rs:11-14`), so the
//! reference's `Particle`, `SolidParticle`, `Material`, `MatInteraction` and `AABB` buffers can be
//! passed by pointer cast; the `const` assertions below pin the sizes and offsets.
//!
//! `seams` shows the three places of the reference's `FluidSim` that call these functions.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

pub mod seams;

pub const ARP_OK: c_int = 0;
pub const ARP_E_ARG: c_int = -1;
pub const ARP_E_HIP: c_int = -2;
pub const ARP_E_STATE: c_int = -3;
pub const ARP_E_NOMEM: c_int = -4;
pub const ARP_E_LIMIT: c_int = -5;
pub const ARP_INVALID_INDEX: u32 = 0xFFFF_FFFF;
pub const ARP_MAX_STAGES: usize = 6;
pub const ARP_VERLET: i32 = -1;
pub const ARP_NL_ALWAYS_BUILD: u32 = 1;
pub const ARP_NL_NO_SPECULATION: u32 = 2;
pub const ARP_BATCH_GRAPH: u32 = 1;

/// `Particle` (particle_state/mod.rs:24-29), std430: 48 bytes.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ArpParticle {
    pub den: f32,
    pub mat: u32,
    pub solid_id: u32,
    pub _pad0: u32,
    pub pos: [f32; 4],
    pub vel: [f32; 4],
}

/// `SolidParticle` (particle_state/mod.rs:31-35), std430: 96 bytes; `stress` is a
/// column-major mat4 holding the plastic strain.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ArpSolidParticle {
    pub part_id: u32,
    pub _pad0: [u32; 3],
    pub ref_pos: [f32; 4],
    pub stress: [f32; 16],
}

/// `Material` (material_region/material.rs:11-28), std430 with bool as 4 bytes: 80 bytes.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ArpMaterial {
    pub immobile: u32,
    pub mass: f32,
    pub start_den: f32,
    pub state_eq: u32,
    pub sound_speed: f32,
    pub target_den: f32,
    pub visc: f32,
    pub bulk_visc: f32,
    pub strain_order: i32,
    pub normal_stiffness: f32,
    pub shear_stiffness: f32,
    pub normal_damp: f32,
    pub shear_damp: f32,
    pub plastic: u32,
    pub yield_strength: f32,
    pub work_hardening: f32,
    pub work_hardening_exp: f32,
    pub kinematic_hardening: f32,
    pub thermal_softening: f32,
    pub relaxation_time: f32,
}

/// `MatInteraction` (material_region/interaction.rs:10-15): 20 bytes.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ArpMatInteraction {
    pub potential: u32,
    pub strength: f32,
    pub radius: f32,
    pub dampening: f32,
    pub friction: f32,
}

/// `AABB` (material_region/region.rs:11-14): 32 bytes.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ArpAABB {
    pub min: [f32; 4],
    pub dim: [f32; 4],
}

/// An explicit Runge-Kutta tableau, or velocity Verlet when `stages == ARP_VERLET`.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ArpTableau {
    pub stages: i32,
    pub a: [[f32; ARP_MAX_STAGES]; ARP_MAX_STAGES],
    pub b: [f32; ARP_MAX_STAGES],
    pub c: [f32; ARP_MAX_STAGES],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ArpGridInfo {
    pub sub: [u32; 4],
    pub ncells: u64,
    pub n_valid: [u32; 3],
    pub n_overflow: [u32; 3],
    pub n_nonfinite: [u32; 3],
    pub occupied_rows: u32,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct ArpSceneSettings {
    pub title: [c_char; 128],
    pub integrator: [c_char; 32],
    pub kernel_radius: f32,
    pub time_step: f32,
    pub artificial_viscosity: f32,
    pub gravity: f32,
    pub render_radius: f32,
    pub subticks: u32,
    pub lighting: u32,
    pub bounds: ArpAABB,
    pub n_objects: u32,
    pub n_on_click: u32,
    pub n_particles: u32,
    pub n_boundary: u32,
    pub n_solids: u32,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct ArpSceneObject {
    pub name: [c_char; 64],
    pub colors: [[f32; 4]; 3],
    pub packing: f32,
    pub immobile: u32,
    pub on_click: u32,
    pub bound: ArpAABB,
    pub material: ArpMaterial,
    pub first: u32,
    pub count: u32,
    pub lattice_points: u32,
}

/// Opaque handles.
#[repr(C)]
pub struct ArpCtx {
    _p: [u8; 0],
}
#[repr(C)]
pub struct ArpBatch {
    _p: [u8; 0],
}
#[repr(C)]
pub struct ArpSceneFile {
    _p: [u8; 0],
}

// The byte contract with include/arp.h (checked against the C compiler's layout by the
// repository's tests, which compile the header and compare these numbers).
macro_rules! offset_of {
    ($t:ty, $f:ident) => {{
        let u = std::mem::MaybeUninit::<$t>::uninit();
        let p = u.as_ptr();
        #[allow(unused_unsafe)]
        unsafe {
            (std::ptr::addr_of!((*p).$f) as usize) - (p as usize)
        }
    }};
}
const _: () = assert!(std::mem::size_of::<ArpParticle>() == 48);
const _: () = assert!(std::mem::size_of::<ArpSolidParticle>() == 96);
const _: () = assert!(std::mem::size_of::<ArpMaterial>() == 80);
const _: () = assert!(std::mem::size_of::<ArpMatInteraction>() == 20);
const _: () = assert!(std::mem::size_of::<ArpAABB>() == 32);
const _: () = assert!(std::mem::size_of::<ArpTableau>() == 196);
const _: () = assert!(std::mem::size_of::<ArpGridInfo>() == 64);
const _: () = assert!(std::mem::size_of::<ArpSceneSettings>() == 240);
const _: () = assert!(std::mem::size_of::<ArpSceneObject>() == 248);

/// Field offsets (run-time checks: offset computation is not `const` on older compilers).
pub fn layout_ok() -> bool {
    offset_of!(ArpParticle, pos) == 16
        && offset_of!(ArpParticle, vel) == 32
        && offset_of!(ArpSolidParticle, ref_pos) == 16
        && offset_of!(ArpSolidParticle, stress) == 32
        && offset_of!(ArpMaterial, strain_order) == 32
        && offset_of!(ArpMaterial, relaxation_time) == 76
        && offset_of!(ArpGridInfo, ncells) == 16
        && offset_of!(ArpGridInfo, n_valid) == 24
        && offset_of!(ArpSceneSettings, kernel_radius) == 160
        && offset_of!(ArpSceneSettings, bounds) == 188
        && offset_of!(ArpSceneSettings, n_objects) == 220
        && offset_of!(ArpSceneObject, packing) == 112
        && offset_of!(ArpSceneObject, bound) == 124
        && offset_of!(ArpSceneObject, material) == 156
        && offset_of!(ArpSceneObject, first) == 236
}

#[link(name = "arp")]
extern "C" {
    pub fn arp_create(device: c_int, flags: u32, out: *mut *mut ArpCtx) -> c_int;
    pub fn arp_destroy(ctx: *mut ArpCtx);
    pub fn arp_last_error(ctx: *const ArpCtx) -> *const c_char;
    pub fn arp_set_stream(ctx: *mut ArpCtx, hip_stream: *mut c_void) -> c_int;
    pub fn arp_synchronize(ctx: *mut ArpCtx) -> c_int;
    pub fn arp_norm_const(dim: u32, h: f32) -> f32;
    pub fn arp_default_interaction(a: *const ArpMaterial, b: *const ArpMaterial, h: f32, out: *mut ArpMatInteraction);
    pub fn arp_resolve_interactions(nm: u32, mats: *const ArpMaterial, cfg: *const ArpMatInteraction, present: *const u8, h: f32, out: *mut ArpMatInteraction);
    pub fn arp_tableau_by_name(name: *const c_char, out: *mut ArpTableau) -> c_int;
    pub fn arp_scene_upload(ctx: *mut ArpCtx, particles: *const ArpParticle, n: u32, solids: *const ArpSolidParticle, ns: u32, boundary: *const ArpParticle, nb: u32, mats: *const ArpMaterial, nm: u32, inter: *const ArpMatInteraction, bounds: *const ArpAABB, kernel_radius: f32, gravity: f32, artificial_viscosity: f32) -> c_int;
    pub fn arp_set_integrator(ctx: *mut ArpCtx, tableau: *const ArpTableau, timestep: f32, subticks: u32) -> c_int;
    pub fn arp_update(ctx: *mut ArpCtx, nframes: u32) -> c_int;
    pub fn arp_step(ctx: *mut ArpCtx, dt: f32, nsteps: u32) -> c_int;
    pub fn arp_time(ctx: *const ArpCtx) -> f64;
    pub fn arp_download(ctx: *mut ArpCtx, particles: *mut ArpParticle, n: u32, solids: *mut ArpSolidParticle, ns: u32) -> c_int;
    pub fn arp_upload_state(ctx: *mut ArpCtx, particles: *const ArpParticle, n: u32, solids: *const ArpSolidParticle, ns: u32) -> c_int;
    pub fn arp_set_neighbor_skin(ctx: *mut ArpCtx, skin: f32, cap: u32, flags: u32) -> c_int;
    pub fn arp_neighbor_lists_fit(cap: u32, n: u32) -> c_int;
    pub fn arp_set_neighbor_max_age(ctx: *mut ArpCtx, stages: u32) -> c_int;
    pub fn arp_neighbor_stats(ctx: *mut ArpCtx, builds: *mut u32, reused: *mut u32) -> c_int;
    pub fn arp_set_neighbor_short_skin(ctx: *mut ArpCtx, short_skin: f32) -> c_int;
    pub fn arp_neighbor_refreshes(ctx: *mut ArpCtx, refreshes: *mut u32) -> c_int;
    pub fn arp_tune_lists(ctx: *mut ArpCtx, dt: f32, trial_steps: u32, skin: *mut f32, short_skin: *mut f32) -> c_int;
    pub fn arp_speculation_stats(ctx: *const ArpCtx, steps: *mut u64, aborted: *mut u64) -> c_int;
    pub fn arp_compute_forces(ctx: *mut ArpCtx, d_state: *const ArpParticle, d_solids: *const ArpSolidParticle, d_deriv: *mut ArpParticle, d_solid_deriv: *mut ArpSolidParticle) -> c_int;
    pub fn arp_lincomb(ctx: *mut ArpCtx, nterms: u32, coefs: *const f32, d_terms: *const *const ArpParticle, d_solid_terms: *const *const ArpSolidParticle, d_out: *mut ArpParticle, d_solid_out: *mut ArpSolidParticle) -> c_int;
    pub fn arp_lincomb_noalloc(ctx: *mut ArpCtx, nterms: u32, coefs: *const f32, d_terms: *const *const ArpParticle, d_solid_terms: *const *const ArpSolidParticle, d_out: *mut ArpParticle, d_solid_out: *mut ArpSolidParticle) -> c_int;
    pub fn arp_velocity(ctx: *mut ArpCtx, r: f32, d_in: *const ArpParticle, d_solid_in: *const ArpSolidParticle, d_out: *mut ArpParticle, d_solid_out: *mut ArpSolidParticle) -> c_int;
    pub fn arp_neighbors_update(ctx: *mut ArpCtx, d_state: *const ArpParticle) -> c_int;
    pub fn arp_grid_info(ctx: *mut ArpCtx, out: *mut ArpGridInfo) -> c_int;
    pub fn arp_neighbor_dump(ctx: *mut ArpCtx, cell: *mut u32, cand_count: *mut u32, cand: *mut u32, cand_capacity: u64) -> c_int;
    pub fn arp_cell_counts(ctx: *mut ArpCtx, first: u64, count: u64, out: *mut u32) -> c_int;
    pub fn arp_profile(ctx: *mut ArpCtx, enable: c_int) -> c_int;
    pub fn arp_profile_query(ctx: *mut ArpCtx, kernel: *const c_char, total_ms: *mut f64, launches: *mut u64) -> c_int;
    pub fn arp_profile_durations(ctx: *mut ArpCtx, kernel: *const c_char, ms: *mut f32, capacity: u64, count: *mut u64) -> c_int;
    pub fn arp_profile_reset(ctx: *mut ArpCtx) -> c_int;
    pub fn arp_profile_select(ctx: *mut ArpCtx, kernels: *const c_char) -> c_int;
    pub fn arp_adjoint_record(ctx: *mut ArpCtx, dt: f32, nsteps: u32, checkpoint_every: u32) -> c_int;
    pub fn arp_adjoint_backward(ctx: *mut ArpCtx, d_grad_out: *const c_void, d_grad_in: *mut c_void, grad_params: *mut f32) -> c_int;
    pub fn arp_set_stiffness_field(ctx: *mut ArpCtx, scale: *const f32, ns: u32) -> c_int;
    pub fn arp_adjoint_stiffness_grad(ctx: *mut ArpCtx, out: *mut f32, ns: u32) -> c_int;
    pub fn arp_batch_create(ctxs: *const *mut ArpCtx, nscenes: u32, flags: u32, out: *mut *mut ArpBatch) -> c_int;
    pub fn arp_batch_destroy(batch: *mut ArpBatch);
    pub fn arp_batch_last_error(batch: *const ArpBatch) -> *const c_char;
    pub fn arp_batch_step(batch: *mut ArpBatch, dt: f32, nsteps: u32) -> c_int;
    pub fn arp_batch_adjoint_record(batch: *mut ArpBatch, dt: f32, nsteps: u32, checkpoint_every: u32) -> c_int;
    pub fn arp_batch_adjoint_backward(batch: *mut ArpBatch, d_grad_out: *const *const c_void, d_grad_in: *const *mut c_void, grad_params: *const *mut f32) -> c_int;
    pub fn arp_batch_tune(batch: *mut ArpBatch, dt: f32, trial_steps: u32, chosen_div: *mut u32) -> c_int;
    pub fn arp_batch_tuning(batch: *const ArpBatch, rowsort_div: *mut u32, walk_buffered: *mut u32) -> c_int;
    pub fn arp_batch_tune_lists(batch: *mut ArpBatch, dt: f32, trial_steps: u32, skin: *mut f32, short_skin: *mut f32) -> c_int;
    pub fn arp_kernel_radius(ctx: *const ArpCtx) -> f32;
    pub fn arp_num_materials(ctx: *const ArpCtx) -> u32;
    pub fn arp_num_particles(ctx: *const ArpCtx) -> u32;
    pub fn arp_num_solids(ctx: *const ArpCtx) -> u32;
    pub fn arp_add_particles(ctx: *mut ArpCtx, material: *const ArpMaterial, particles: *const ArpParticle, count: u32, material_added: *mut u32) -> c_int;
    pub fn arp_scene_parse(toml_text: *const c_char, len: usize, base_dir: *const c_char, out: *mut *mut ArpSceneFile, err: *mut c_char, err_len: usize) -> c_int;
    pub fn arp_scene_free(f: *mut ArpSceneFile);
    pub fn arp_scene_error(f: *const ArpSceneFile) -> *const c_char;
    pub fn arp_scene_settings(f: *const ArpSceneFile, out: *mut ArpSceneSettings) -> c_int;
    pub fn arp_scene_object(f: *const ArpSceneFile, i: u32, out: *mut ArpSceneObject) -> c_int;
    pub fn arp_scene_generate(f: *mut ArpSceneFile, threads: u32) -> c_int;
    pub fn arp_scene_arrays(f: *const ArpSceneFile, particles: *mut ArpParticle, n: u32, solids: *mut ArpSolidParticle, ns: u32, boundary: *mut ArpParticle, nb: u32, mats: *mut ArpMaterial, nm: u32, inter: *mut ArpMatInteraction) -> c_int;
    pub fn arp_scene_contains(f: *const ArpSceneFile, object: u32, points: *const f32, n: u32, out: *mut u8) -> c_int;
    pub fn arp_scene_apply(ctx: *mut ArpCtx, f: *const ArpSceneFile) -> c_int;
    pub fn arp_scene_spawn(ctx: *mut ArpCtx, f: *const ArpSceneFile, object: u32, offset: *const f32, material_added: *mut u32) -> c_int;
    pub fn arp_toml_to_json(text: *const c_char, len: usize, out: *mut c_char, cap: usize, needed: *mut usize) -> c_int;
}

/// The message of the last failed call on `ctx`.
pub fn last_error(ctx: *const ArpCtx) -> String {
    unsafe {
        let p = arp_last_error(ctx);
        if p.is_null() {
            String::new()
        } else {
            std::ffi::CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    }
}

#[cfg(test)]
mod tests {
    #[test]
    fn layout() {
        assert!(super::layout_ok());
    }
}
