//! ae_mpm_hip -- AE-ENGINE's MLS-MPM step on an MI355X (gfx950) through the
//! C ABI of libaemh.so (`include/aemh.h`).
//!
//! `ffi` declares every entry point of the header, in the pattern of
//! ae_cpp_kernel's `extern "C"` block (ae_cpp_kernel/src/lib.rs:21-32).
//! `MpmGpu::mpm_step_parallel_with_grid` keeps the call shape of
//! `ae_particle::mpm_solver::mpm_step_parallel_with_grid`
//! (ae_particle/src/mpm_solver.rs:1063-1084), so the engine's call site
//! (ae_engine/src/managers/simulation.rs:573-581) swaps one call for the
//! other.

use std::ffi::CStr;
use std::os::raw::{c_char, c_int};

use ae_particle::constitutive::{ConstitutiveModel, MaterialParams};
use ae_particle::mpm_solver::{Grid, MpmConfig};
use ae_particle::mpss::MpssBuffer;

pub mod ffi {
    use std::os::raw::{c_char, c_int};

    pub const AEMH_OK: c_int = 0;
    pub const AEMH_ERR_ARG: c_int = 1;
    pub const AEMH_ERR_HIP: c_int = 2;
    pub const AEMH_ERR_OOM: c_int = 3;
    pub const AEMH_ERR_STATE: c_int = 4;
    pub const AEMH_ERR_OVERFLOW: c_int = 5;
    pub const AEMH_ERR_NO_DEVICE: c_int = 6;

    pub const AEMH_ACCUM_F32: u32 = 0;
    pub const AEMH_ACCUM_Q32_32: u32 = 1;
    pub const AEMH_ACCUM_F32_SERIAL: u32 = 2;
    pub const AEMH_CLAMP_AUTO: u32 = 0;
    pub const AEMH_CLAMP_SERIAL: u32 = 1;
    pub const AEMH_CLAMP_WINDOW: u32 = 2;
    pub const AEMH_STORE_F32: u32 = 0;
    pub const AEMH_STORE_BF16: u32 = 1;
    pub const AEMH_STORE_F16: u32 = 2;
    pub const AEMH_NO_SLOT: u32 = 0xFFFF_FFFF;

    #[repr(C)]
    pub struct AemhCtx {
        _p: [u8; 0],
    }
    #[repr(C)]
    pub struct AemhFixed {
        _p: [u8; 0],
    }

    #[repr(C)]
    #[derive(Clone, Copy, Default, Debug)]
    pub struct AemhConfig {
        pub substeps: u32,
        pub gravity: [f32; 3],
        pub damping: f32,
        pub bc_margin: u32,
        pub bc_restitution: f32,
        pub accum: u32,
        pub clamp_mode: u32,
        pub storage: u32,
        pub resort_interval: u32,
        pub reserved: [u32; 3],
    }

    #[repr(C)]
    #[derive(Clone, Copy, Default, Debug)]
    pub struct AemhScene {
        pub first: u32,
        pub count: u32,
        pub nx: u32,
        pub ny: u32,
        pub nz: u32,
        pub dx: f32,
        pub origin: [f32; 3],
        pub dt: f32,
        pub velocity_only: u32,
        pub clamp_mode: u32,
        pub reserved: u32,
    }

    #[repr(C)]
    #[derive(Clone, Copy, Default, Debug)]
    pub struct AemhLodConfig {
        pub near_field_distance: f32,
        pub mid_field_distance: f32,
        pub far_field_update_interval: u32,
        pub max_near_particles: u32,
        pub mid_interval: u32,
        pub near_n: u32,
        pub near_dx: f32,
        pub mid_n: u32,
        pub mid_dx: f32,
        pub reserved: [u32; 3],
    }

    #[repr(C)]
    #[derive(Clone, Copy, Default, Debug)]
    pub struct AemhTierStats {
        pub near: u32,
        pub mid: u32,
        pub far: u32,
        pub reclassified: u32,
        pub near_ran: u32,
        pub mid_ran: u32,
        pub far_ran: u32,
        pub expired: u32,
        pub near_origin: [f32; 3],
        pub mid_origin: [f32; 3],
    }

    #[repr(C)]
    #[derive(Clone, Copy, Default, Debug)]
    pub struct AemhFixedConfig {
        pub grid_resolution: [u32; 3],
        pub cell_size: i64,
        pub youngs_modulus: i64,
        pub poissons_ratio: i64,
        pub yield_stress: i64,
        pub hardening: i64,
        pub density: i64,
        pub gravity: [i64; 3],
        pub dt: i64,
        pub substeps: u32,
        /// MpmMaterialModel order (ae_physics/src/mpm.rs:5-13): 0 Elastic .. 5 Snow
        pub material_model: u32,
        /// friction_angle | fracture_strain | viscosity | hardening, as raw Q32.32
        pub model_param: i64,
        pub enable_fracture: u32,
        pub fracture_strain: i64,
        pub particle_count: u64,
    }

    #[repr(C)]
    #[derive(Clone, Copy, Default, Debug)]
    pub struct AemhFixedState {
        pub particles: u64,
        pub step_count: u64,
        pub time: f64,
        pub next_particle_id: u64,
        pub fracture_events: u64,
        pub hardening: i64,
        pub fracture_strain: i64,
        pub enable_fracture: u32,
        pub material_model: u32,
        pub model_param: i64,
        pub checked_substeps: u64,
        pub serial_substeps: u64,
        pub max_scatter_bound: u64,
    }

    #[repr(C)]
    #[derive(Clone, Copy, Default, Debug)]
    pub struct AemhFractureEvent {
        pub time: f64,
        pub position: [f32; 3],
        pub strain_magnitude: f32,
        pub num_fragments: u32,
        pub pad: u32,
    }

    #[repr(C)]
    #[derive(Clone, Copy, Default, Debug)]
    pub struct AemhFixedStats {
        pub active_particles: u64,
        pub fractured_count: u64,
        pub total_mass: i64,
        pub average_temperature: i64,
        pub kinetic_energy: i64,
        pub total_strain_energy: i64,
    }

    #[repr(C)]
    #[derive(Clone, Copy, Default, Debug)]
    pub struct AemhFixedSnapshot {
        pub position: [f32; 3],
        pub velocity: [f32; 3],
        pub stress_magnitude: f32,
        pub plastic_strain: f32,
        pub temperature: f32,
        pub color: [f32; 4],
        pub fractured: u32,
    }

    extern "C" {
        pub fn aemh_config_default(cfg: *mut AemhConfig);
        pub fn aemh_version() -> *const c_char;

        pub fn aemh_create(out: *mut *mut AemhCtx, device: c_int) -> c_int;
        pub fn aemh_destroy(ctx: *mut AemhCtx) -> c_int;
        pub fn aemh_last_error(ctx: *const AemhCtx) -> *const c_char;

        pub fn aemh_set_config(ctx: *mut AemhCtx, cfg: *const AemhConfig) -> c_int;
        pub fn aemh_set_grid(ctx: *mut AemhCtx, nx: u32, ny: u32, nz: u32, dx: f32,
                             origin: *const f32) -> c_int;
        pub fn aemh_set_materials(ctx: *mut AemhCtx, n: u32, params7: *const f32,
                                  models: *const u8, ext3: *const f32) -> c_int;

        pub fn aemh_upload_particles(ctx: *mut AemhCtx, n: u32, pos3: *const f32,
                                     vel3: *const f32, c9: *const f32, f9: *const f32,
                                     j: *const f32, mass: *const f32, mat: *const u16) -> c_int;
        pub fn aemh_upload_plastic_j(ctx: *mut AemhCtx, jp: *const f32) -> c_int;
        pub fn aemh_download_particles(ctx: *mut AemhCtx, pos3: *mut f32, vel3: *mut f32,
                                       c9: *mut f32, f9: *mut f32, j: *mut f32) -> c_int;
        pub fn aemh_download_plastic_j(ctx: *mut AemhCtx, jp: *mut f32) -> c_int;
        pub fn aemh_set_indices(ctx: *mut AemhCtx, idx: *const u32, n_idx: u32) -> c_int;

        pub fn aemh_mpss_reserve(ctx: *mut AemhCtx, capacity: u32) -> c_int;
        pub fn aemh_mpss_spawn(ctx: *mut AemhCtx, n: u32, slots_out: *mut u32,
                               n_spawned: *mut u32) -> c_int;
        pub fn aemh_mpss_kill(ctx: *mut AemhCtx, slots: *const u32, n: u32) -> c_int;
        pub fn aemh_mpss_compact(ctx: *mut AemhCtx) -> c_int;
        pub fn aemh_mpss_update_lifetimes(ctx: *mut AemhCtx, dt: f32, expired_out: *mut u32,
                                          max_out: u32, n_expired: *mut u32) -> c_int;
        pub fn aemh_mpss_write(ctx: *mut AemhCtx, slots: *const u32, n: u32, pos3: *const f32,
                               vel3: *const f32, c9: *const f32, f9: *const f32, j: *const f32,
                               jp: *const f32, mass: *const f32, mat: *const u16,
                               lifetime: *const f32, age: *const f32) -> c_int;
        pub fn aemh_mpss_read(ctx: *mut AemhCtx, slots: *const u32, n: u32, pos3: *mut f32,
                              vel3: *mut f32, c9: *mut f32, f9: *mut f32, j: *mut f32,
                              jp: *mut f32, mass: *mut f32, mat: *mut u16, lifetime: *mut f32,
                              age: *mut f32) -> c_int;
        pub fn aemh_mpss_state(ctx: *mut AemhCtx, count: *mut u32, capacity: *mut u32,
                               active: *mut u8) -> c_int;

        pub fn aemh_lod_config_default(lod: *mut AemhLodConfig);
        pub fn aemh_tier_set_config(ctx: *mut AemhCtx, lod: *const AemhLodConfig) -> c_int;
        pub fn aemh_tier_tick(ctx: *mut AemhCtx, player: *const f32, dt: f32,
                              stats: *mut AemhTierStats) -> c_int;
        pub fn aemh_tier_lists(ctx: *mut AemhCtx, near: *mut u32, mid: *mut u32, far: *mut u32,
                               counts3: *mut u32) -> c_int;
        pub fn aemh_tier_render(ctx: *mut AemhCtx, tier: u32, take: u32, pos3: *mut f32,
                                slots: *mut u32, n_out: *mut u32) -> c_int;

        pub fn aemh_step(ctx: *mut AemhCtx, total_dt: f32) -> c_int;
        pub fn aemh_mpm_step_parallel_with_grid(ctx: *mut AemhCtx, idx: *const u32, n_idx: u32,
                                                total_dt: f32) -> c_int;
        pub fn aemh_step_velocity_only(ctx: *mut AemhCtx, total_dt: f32) -> c_int;
        pub fn aemh_step_async(ctx: *mut AemhCtx, total_dt: f32) -> c_int;
        pub fn aemh_sync(ctx: *mut AemhCtx) -> c_int;
        pub fn aemh_set_scenes(ctx: *mut AemhCtx, scenes: *const AemhScene, n: u32) -> c_int;
        pub fn aemh_batch_step(ctx: *mut AemhCtx, n_substeps: u32) -> c_int;
        pub fn aemh_autotune(ctx: *mut AemhCtx) -> c_int;
        pub fn aemh_get_launch_shape(ctx: *mut AemhCtx, p2g_tiles: *mut u32,
                                     p2g_partials: *mut u32, g2p_wg: *mut u32,
                                     upd_wgs: *mut u32, tuned_ms3: *mut f32) -> c_int;
        // p2g_partials: bit 0 partial-sum tile flush, bit 1 pair reduction (FP32 mode),
        // bit 2 lane layers
        pub fn aemh_set_launch_shape(ctx: *mut AemhCtx, p2g_tiles: u32, p2g_partials: u32,
                                     g2p_wg: u32, upd_wgs: u32) -> c_int;

        pub fn aemh_download_grid(ctx: *mut AemhCtx, scene: u32, mass: *mut f32,
                                  momentum3: *mut f32, force3: *mut f32, vel3: *mut f32) -> c_int;
        pub fn aemh_download_grid_q32(ctx: *mut AemhCtx, scene: u32, q7: *mut i64) -> c_int;
        pub fn aemh_conservation_f64(ctx: *mut AemhCtx, out5: *mut f64) -> c_int;
        pub fn aemh_conservation_f64_noalloc(ctx: *mut AemhCtx, out5: *mut f64) -> c_int;
        pub fn aemh_svd_fp64_count(ctx: *mut AemhCtx, count: *mut u64) -> c_int;
        pub fn aemh_last_substep_info(ctx: *mut AemhCtx, active_blocks: *mut u32,
                                      particles: *mut u32) -> c_int;
        pub fn aemh_profile_enable(ctx: *mut AemhCtx, on: c_int) -> c_int;
        pub fn aemh_profile_get(ctx: *mut AemhCtx, kernel: *const c_char, avg_ms: *mut f64,
                                count: *mut u64) -> c_int;
        pub fn aemh_profile_reset(ctx: *mut AemhCtx) -> c_int;
        pub fn aemh_scene_uniform_box(seed: u64, n: u32, lo: *const f32, hi: *const f32,
                                      pos3: *mut f32);

        pub fn aemh_fixed_config_default(cfg: *mut AemhFixedConfig);
        pub fn aemh_fixed_create(out: *mut *mut AemhFixed, device: c_int,
                                 cfg: *const AemhFixedConfig) -> c_int;
        pub fn aemh_fixed_destroy(sim: *mut AemhFixed);
        pub fn aemh_fixed_last_error(sim: *const AemhFixed) -> *const c_char;
        pub fn aemh_fixed_add_particle(sim: *mut AemhFixed, pos3: *const i64, vel3: *const i64,
                                       mass: i64, color4: *const f32, id_out: *mut u64) -> c_int;
        pub fn aemh_fixed_add_particles(sim: *mut AemhFixed, n: u64, pos3: *const i64,
                                        vel3: *const i64, mass: *const i64, color4: *const f32,
                                        first_id_out: *mut u64) -> c_int;
        pub fn aemh_fixed_remove_particle(sim: *mut AemhFixed, id: u64, removed: *mut c_int)
                                          -> c_int;
        pub fn aemh_fixed_step(sim: *mut AemhFixed, n_steps: u32) -> c_int;
        pub fn aemh_fixed_set_material_model(sim: *mut AemhFixed, model: u32, model_param: i64)
                                             -> c_int;
        pub fn aemh_fixed_set_sort_interval(sim: *mut AemhFixed, substeps: u32) -> c_int;
        pub fn aemh_fixed_enable_fracture(sim: *mut AemhFixed, threshold: i64) -> c_int;
        pub fn aemh_fixed_disable_fracture(sim: *mut AemhFixed) -> c_int;
        pub fn aemh_fixed_get_state(sim: *mut AemhFixed, out: *mut AemhFixedState) -> c_int;
        pub fn aemh_fixed_read_particles(sim: *mut AemhFixed, first: u64, count: u64,
                                         q29: *mut i64, flags3: *mut u8, ids9: *mut u64,
                                         color4: *mut f32) -> c_int;
        pub fn aemh_fixed_write_particles(sim: *mut AemhFixed, first: u64, count: u64,
                                          q29: *const i64, flags3: *const u8, ids9: *const u64,
                                          color4: *const f32) -> c_int;
        pub fn aemh_fixed_fracture_events(sim: *mut AemhFixed, first: u64, count: u64,
                                          out: *mut AemhFractureEvent) -> c_int;
        pub fn aemh_fixed_get_stats(sim: *mut AemhFixed, out: *mut AemhFixedStats) -> c_int;
        pub fn aemh_fixed_snapshots(sim: *mut AemhFixed, out: *mut AemhFixedSnapshot, cap: u64,
                                    n_out: *mut u64) -> c_int;
    }
}

use ffi::*;

#[derive(Debug)]
pub struct AemhError {
    pub code: i32,
    pub message: String,
}

/// One GPU context stepping the engine's `MpssBuffer` call by call.
///
/// The host buffer stays the owner of the particles, as in the reference:
/// every call uploads the slots `particle_indices` names (the only slots the
/// step reads or writes, mpm_solver.rs:1088-1326) from the buffer as the
/// caller left it -- after spawns, field writes or the engine's
/// `apply_boundary_conditions_par` (ae_engine/src/lib.rs:346-349) -- and
/// writes them back after the step.  Nothing of a previous call is trusted.
/// The context holds device copies of `capacity` slots (aemh_mpss_reserve),
/// so slots cross PCIe only when listed.
pub struct MpmGpu {
    ctx: *mut AemhCtx,
    capacity: usize,
    // host staging rows of the listed slots, reused between calls
    idx: Vec<u32>,
    pos: Vec<f32>,
    vel: Vec<f32>,
    c: Vec<f32>,
    f: Vec<f32>,
    j: Vec<f32>,
    mass: Vec<f32>,
    mat: Vec<u16>,
}

impl MpmGpu {
    pub fn new(device: i32) -> Result<Self, AemhError> {
        let mut ctx = std::ptr::null_mut();
        let rc = unsafe { aemh_create(&mut ctx, device) };
        if rc != AEMH_OK {
            return Err(AemhError { code: rc, message: "aemh_create".into() });
        }
        Ok(Self {
            ctx,
            capacity: 0,
            idx: Vec::new(),
            pos: Vec::new(),
            vel: Vec::new(),
            c: Vec::new(),
            f: Vec::new(),
            j: Vec::new(),
            mass: Vec::new(),
            mat: Vec::new(),
        })
    }

    pub fn ctx(&self) -> *mut AemhCtx {
        self.ctx
    }

    pub fn check(&self, rc: c_int) -> Result<(), AemhError> {
        if rc == AEMH_OK {
            return Ok(());
        }
        let msg = unsafe { CStr::from_ptr(aemh_last_error(self.ctx) as *const c_char) };
        Err(AemhError { code: rc, message: msg.to_string_lossy().into_owned() })
    }

    /// Same arguments and meaning as
    /// `ae_particle::mpm_solver::mpm_step_parallel_with_grid`
    /// (mpm_solver.rs:1063-1084).  The grid's contents stay on the device
    /// (`aemh_download_grid` reads them); the buffer's listed slots are
    /// updated in place and `update_lifetimes(total_dt)` runs on the host
    /// buffer, as the reference's last line does (:1083).
    #[allow(clippy::too_many_arguments)]
    pub fn mpm_step_parallel_with_grid(
        &mut self,
        buffer: &mut MpssBuffer,
        grid: &Grid,
        config: &MpmConfig,
        material_table: &[MaterialParams],
        model_table: &[ConstitutiveModel],
        total_dt: f32,
        particle_indices: &[usize],
    ) -> Result<(), AemhError> {
        if particle_indices.is_empty() {
            return Ok(()); // mpm_solver.rs:1072-1074
        }
        let mut cfg = AemhConfig::default();
        unsafe { aemh_config_default(&mut cfg) };
        cfg.substeps = config.substeps as u32;
        cfg.gravity = config.gravity;
        cfg.damping = config.damping;
        // below 50,000 indices the reference sums P2G serially in list order
        // (mpm_solver.rs:1077-1081): the reference-order mode reproduces it
        if particle_indices.len() < 50_000 {
            cfg.accum = AEMH_ACCUM_F32_SERIAL;
        }
        self.check(unsafe { aemh_set_config(self.ctx, &cfg) })?;
        self.check(unsafe {
            aemh_set_grid(self.ctx, grid.nx as u32, grid.ny as u32, grid.nz as u32, grid.dx,
                          grid.origin.as_ptr())
        })?;
        self.set_materials(material_table, model_table)?;
        if self.capacity != buffer.capacity {
            self.check(unsafe { aemh_mpss_reserve(self.ctx, buffer.capacity as u32) })?;
            self.capacity = buffer.capacity;
        }
        self.gather(buffer, particle_indices);
        let n = self.idx.len() as u32;
        self.check(unsafe {
            aemh_mpss_write(self.ctx, self.idx.as_ptr(), n, self.pos.as_ptr(), self.vel.as_ptr(),
                            self.c.as_ptr(), self.f.as_ptr(), self.j.as_ptr(), std::ptr::null(),
                            self.mass.as_ptr(), self.mat.as_ptr(), std::ptr::null(),
                            std::ptr::null())
        })?;
        self.check(unsafe {
            aemh_mpm_step_parallel_with_grid(self.ctx, self.idx.as_ptr(), n, total_dt)
        })?;
        self.check(unsafe {
            aemh_mpss_read(self.ctx, self.idx.as_ptr(), n, self.pos.as_mut_ptr(),
                           self.vel.as_mut_ptr(), self.c.as_mut_ptr(), self.f.as_mut_ptr(),
                           self.j.as_mut_ptr(), std::ptr::null_mut(), std::ptr::null_mut(),
                           std::ptr::null_mut(), std::ptr::null_mut(), std::ptr::null_mut())
        })?;
        self.scatter(buffer);
        let _ = buffer.update_lifetimes(total_dt);
        Ok(())
    }

    /// material_table / model_table as the reference looks them up
    /// (mpm_solver.rs:839-841): index i takes material_table.get(i), else
    /// entry 0, and model_table.get(i), else NeoHookean.  One entry per index
    /// below the longer table; indices past both take entry 0 as NeoHookean,
    /// the library's own fallback.
    fn set_materials(&self, material_table: &[MaterialParams],
                     model_table: &[ConstitutiveModel]) -> Result<(), AemhError> {
        if material_table.is_empty() {
            return Err(AemhError { code: AEMH_ERR_ARG, message: "empty material_table".into() });
        }
        let n = material_table.len().max(model_table.len());
        let mut params = Vec::with_capacity(n * 7);
        let mut models = Vec::with_capacity(n);
        for i in 0..n {
            let m = material_table.get(i).unwrap_or(&material_table[0]);
            params.extend_from_slice(&[m.young_modulus, m.poisson_ratio, m.yield_stress,
                                       m.hardening, m.density, m.friction_angle, m.cohesion]);
            models.push(*model_table.get(i).unwrap_or(&ConstitutiveModel::NeoHookean) as u8);
        }
        self.check(unsafe {
            aemh_set_materials(self.ctx, n as u32, params.as_ptr(), models.as_ptr(),
                               std::ptr::null())
        })
    }

    fn gather(&mut self, b: &MpssBuffer, particle_indices: &[usize]) {
        self.idx.clear();
        self.pos.clear();
        self.vel.clear();
        self.c.clear();
        self.f.clear();
        self.j.clear();
        self.mass.clear();
        self.mat.clear();
        for &i in particle_indices {
            self.idx.push(i as u32);
            self.pos.extend_from_slice(&b.pos[i]);
            self.vel.extend_from_slice(&b.vel[i]);
            self.c.extend_from_slice(&b.c[i]);
            self.f.extend_from_slice(&b.strain[i]);
            self.j.push(b.jacobian[i]);
            self.mass.push(b.mass[i]);
            self.mat.push(b.material_idx[i]);
        }
    }

    fn scatter(&self, b: &mut MpssBuffer) {
        for (k, &i) in self.idx.iter().enumerate() {
            let i = i as usize;
            b.pos[i].copy_from_slice(&self.pos[3 * k..3 * k + 3]);
            b.vel[i].copy_from_slice(&self.vel[3 * k..3 * k + 3]);
            b.c[i].copy_from_slice(&self.c[9 * k..9 * k + 9]);
            b.strain[i].copy_from_slice(&self.f[9 * k..9 * k + 9]);
            b.jacobian[i] = self.j[k];
        }
    }
}

impl Drop for MpmGpu {
    fn drop(&mut self) {
        unsafe {
            aemh_destroy(self.ctx);
        }
    }
}
