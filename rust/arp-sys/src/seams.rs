//! The three seams of the reference's `FluidSim` (`src/soft_body/mod.rs:23-275`) and the
//! per-stage closures (`src/soft_body/shaders/mod.rs:11-86`,
//! `src/soft_body/particle_state/particles.rs:366-446`) that call `libarp.so` instead of GL.
//!
//! `Sim` owns one library context (the device buffers, the HIP stream and the code object);
//! dropping it destroys the context.  The reference keeps its own types: its `Particle`,
//! `SolidParticle`, `Material`, `MatInteraction` and `AABB` slices are passed with
//! `as_wire` (a pointer cast, the layouts being byte-identical).
//!
//! Patched `FluidSim` (the fused path):
//!
//! ```ignore
//! // FluidSim::new, after the interaction matrix (mod.rs:93-111):
//! let sim = arp_sys::seams::Sim::new(0, as_wire(&particles), as_wire(&solids), as_wire(&boundary),
//!     as_wire(&materials), as_wire(&interactions), &as_wire(std::slice::from_ref(&bounds))[0], h, g, f,
//!     &integrator_name, timestep, subticks)?;
//! // FluidSim::update (mod.rs:235-274): the subtick loop and both closures collapse to
//! self.sim.update(1)?;
//! // FluidSim::particles: the state in the caller's order
//! let (p, s) = self.sim.download()?;
//! ```
//!
//! Per-stage drop-ins (the reference keeps `numerical-integration` and `ParticleState`, and
//! one device buffer per term): `compute_forces` replaces the force closure's GL dispatches,
//! `neighbors_update` replaces `NeighborList::update_contents`, `lincomb` replaces
//! `ArithShaders::linear_combination` and `velocity` replaces `ArithShaders::velocity`.

use super::*;
use std::ffi::CString;
use std::ptr;

pub type Result<T> = std::result::Result<T, String>;

/// Reinterprets a slice of a byte-identical wire type (the reference's own std430 structs).
///
/// # Safety
/// `T` must have the size and layout of `W` (checked for size here).
pub unsafe fn as_wire<T, W>(s: &[T]) -> &[W] {
    assert_eq!(std::mem::size_of::<T>(), std::mem::size_of::<W>());
    std::slice::from_raw_parts(s.as_ptr() as *const W, s.len())
}

fn check(ctx: *const ArpCtx, rc: c_int) -> Result<()> {
    if rc == ARP_OK {
        Ok(())
    } else {
        Err(format!("arp error {}: {}", rc, last_error(ctx)))
    }
}

pub struct Sim {
    ctx: *mut ArpCtx,
}

impl Sim {
    /// `FluidSim::new` (soft_body/mod.rs:58-166) + the integrator of examples/test.rs:160-205.
    #[allow(clippy::too_many_arguments)]
    pub fn new(device: i32, particles: &[ArpParticle], solids: &[ArpSolidParticle], boundary: &[ArpParticle],
               materials: &[ArpMaterial], interactions: &[ArpMatInteraction], bounds: &ArpAABB,
               kernel_radius: f32, gravity: f32, artificial_viscosity: f32, integrator: &str,
               timestep: f32, subticks: u32) -> Result<Sim> {
        let nm = materials.len();
        if interactions.len() != nm * nm {
            return Err(format!("interaction matrix has {} entries, {} materials", interactions.len(), nm));
        }
        let mut ctx = ptr::null_mut();
        let rc = unsafe { arp_create(device, 0, &mut ctx) };
        if rc != ARP_OK {
            return Err(format!("arp_create failed ({})", rc));
        }
        let sim = Sim { ctx };
        check(ctx, unsafe {
            arp_scene_upload(ctx, particles.as_ptr(), particles.len() as u32, solids.as_ptr(), solids.len() as u32,
                             boundary.as_ptr(), boundary.len() as u32, materials.as_ptr(), nm as u32,
                             interactions.as_ptr(), bounds, kernel_radius, gravity, artificial_viscosity)
        })?;
        let mut tab = ArpTableau::default();
        let name = CString::new(integrator).map_err(|e| e.to_string())?;
        check(ctx, unsafe { arp_tableau_by_name(name.as_ptr(), &mut tab) })?;
        check(ctx, unsafe { arp_set_integrator(ctx, &tab, timestep, subticks) })?;
        Ok(sim)
    }

    pub fn raw(&self) -> *mut ArpCtx {
        self.ctx
    }

    /// `FluidSim::update` (soft_body/mod.rs:235-274): `frames` frames of `subticks` steps.
    pub fn update(&mut self, frames: u32) -> Result<()> {
        check(self.ctx, unsafe { arp_update(self.ctx, frames) })
    }

    /// `FluidSim::time`.
    pub fn time(&self) -> f64 {
        unsafe { arp_time(self.ctx) }
    }

    /// `FluidSim::particles` and the solid buffer, in the caller's original order.
    pub fn download(&mut self) -> Result<(Vec<ArpParticle>, Vec<ArpSolidParticle>)> {
        let n = unsafe { arp_num_particles(self.ctx) } as usize;
        let ns = unsafe { arp_num_solids(self.ctx) } as usize;
        let mut p = vec![ArpParticle::default(); n];
        let mut s = vec![ArpSolidParticle::default(); ns];
        check(self.ctx, unsafe { arp_download(self.ctx, p.as_mut_ptr(), n as u32, s.as_mut_ptr(), ns as u32) })?;
        Ok((p, s))
    }

    /// `FluidSim::add_particles` (soft_body/mod.rs:180-206): returns whether a material was added.
    pub fn add_particles(&mut self, material: &ArpMaterial, particles: &[ArpParticle]) -> Result<bool> {
        let mut added = 0u32;
        check(self.ctx, unsafe {
            arp_add_particles(self.ctx, material, particles.as_ptr(), particles.len() as u32, &mut added)
        })?;
        Ok(added != 0)
    }

    /// `FluidSim::kernel_radius`, `FluidSim::num_materials`.
    pub fn kernel_radius(&self) -> f32 {
        unsafe { arp_kernel_radius(self.ctx) }
    }
    pub fn num_materials(&self) -> u32 {
        unsafe { arp_num_materials(self.ctx) }
    }

    // ---- per-stage drop-ins on device buffers (raw HIP device pointers) ----------------

    /// The force closure (shaders/mod.rs:11-86): derivative of `d_state` into `d_deriv`.
    ///
    /// # Safety
    /// The pointers are device arrays of the scene's particle / solid counts.
    pub unsafe fn compute_forces(&mut self, d_state: *const ArpParticle, d_solids: *const ArpSolidParticle,
                                 d_deriv: *mut ArpParticle, d_solid_deriv: *mut ArpSolidParticle) -> Result<()> {
        check(self.ctx, arp_compute_forces(self.ctx, d_state, d_solids, d_deriv, d_solid_deriv))
    }

    /// `NeighborList::update_contents` (neighbor_list.rs:224-294).
    ///
    /// # Safety
    /// `d_state` is a device array of the scene's particle count.
    pub unsafe fn neighbors_update(&mut self, d_state: *const ArpParticle) -> Result<()> {
        check(self.ctx, arp_neighbors_update(self.ctx, d_state))
    }

    /// `ArithShaders::linear_combination` (particles.rs:381-436).
    ///
    /// # Safety
    /// Every term and output is a device array of the scene's counts.
    pub unsafe fn lincomb(&mut self, coefs: &[f32], terms: &[*const ArpParticle],
                          solid_terms: Option<&[*const ArpSolidParticle]>, out: *mut ArpParticle,
                          solid_out: *mut ArpSolidParticle) -> Result<()> {
        if coefs.len() != terms.len() || solid_terms.map_or(false, |s| s.len() != terms.len()) {
            return Err("lincomb: one coefficient per term".to_string());
        }
        check(self.ctx, arp_lincomb(self.ctx, terms.len() as u32, coefs.as_ptr(), terms.as_ptr(),
                                    solid_terms.map_or(ptr::null(), |s| s.as_ptr()), out, solid_out))
    }

    /// `ArithShaders::velocity` (particles.rs:438-446).
    ///
    /// # Safety
    /// The pointers are device arrays of the scene's counts.
    pub unsafe fn velocity(&mut self, r: f32, d_in: *const ArpParticle, d_solid_in: *const ArpSolidParticle,
                           d_out: *mut ArpParticle, d_solid_out: *mut ArpSolidParticle) -> Result<()> {
        check(self.ctx, arp_velocity(self.ctx, r, d_in, d_solid_in, d_out, d_solid_out))
    }
}

impl Drop for Sim {
    fn drop(&mut self) {
        unsafe { arp_destroy(self.ctx) }
    }
}
