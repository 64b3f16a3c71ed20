//! The C ABI of libbendy2d.so (include/bendy2d.h): records field for field and
//! the entry points the facade uses.

use std::os::raw::c_char;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct BdVec2 {
    pub x: f32,
    pub y: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct BdParticle {
    pub pos: BdVec2,
    pub prev_pos: BdVec2,
    pub acc: BdVec2,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct BdLink {
    pub particle_a: u64,
    pub particle_b: u64,
    pub target_distance: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct BdCircle {
    pub point: BdParticle,
    pub radius: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct BdBounds {
    pub pos: BdVec2,
    pub size: BdVec2,
}

#[repr(C)]
pub struct BdPolygonIn {
    pub particles: *const BdParticle,
    pub n_particles: u64,
    pub particle_links: *const BdLink,
    pub n_links: u64,
    pub is_static: i32,
    pub center: BdVec2,
    pub scale: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct BdPolygonInfo {
    pub n_particles: u64,
    pub n_links: u64,
    pub is_static: i32,
    pub center: BdVec2,
    pub scale: f32,
}

#[repr(C)]
pub struct BdSolver {
    _opaque: [u8; 0],
}

extern "C" {
    pub fn bd_last_error() -> *const c_char;

    pub fn bd_solver_new() -> *mut BdSolver;
    pub fn bd_solver_free(s: *mut BdSolver);
    pub fn bd_solver_clone(s: *const BdSolver) -> *mut BdSolver;

    pub fn bd_set_gravity(s: *mut BdSolver, g: BdVec2);
    pub fn bd_set_bounds(s: *mut BdSolver, b: BdBounds);
    pub fn bd_set_bounds_active(s: *mut BdSolver, active: i32);

    pub fn bd_add_particle(s: *mut BdSolver, pos: BdVec2) -> i64;
    pub fn bd_add_circle(s: *mut BdSolver, c: *const BdCircle) -> i64;
    pub fn bd_add_polygon(s: *mut BdSolver, p: *const BdPolygonIn) -> i64;
    pub fn bd_add_particle_link(s: *mut BdSolver, l: *const BdLink) -> i64;
    pub fn bd_add_circle_link(s: *mut BdSolver, l: *const BdLink) -> i64;

    pub fn bd_get_particle_len(s: *const BdSolver) -> u64;
    pub fn bd_get_circles_len(s: *const BdSolver) -> u64;
    pub fn bd_get_polygons_len(s: *const BdSolver) -> u64;
    pub fn bd_get_particles(s: *const BdSolver, out: *mut BdParticle, cap: u64) -> u64;
    pub fn bd_get_circles(s: *const BdSolver, out: *mut BdCircle, cap: u64) -> u64;
    pub fn bd_get_polygon(s: *const BdSolver, index: u64, info: *mut BdPolygonInfo) -> i32;
    pub fn bd_get_polygon_particles(s: *const BdSolver, index: u64, out: *mut BdParticle, cap: u64) -> u64;
    pub fn bd_get_polygon_links(s: *const BdSolver, index: u64, out: *mut BdLink, cap: u64) -> u64;

    pub fn bd_update(s: *mut BdSolver, dt: f32) -> i32;
    pub fn bd_ext_set_readback(s: *mut BdSolver, on: i32) -> i32;

    pub fn bd_polygon_circle(radius: f32, pos: BdVec2, point_count: u64, is_static: i32,
                             out_particles: *mut BdParticle, out_links: *mut BdLink,
                             out_center: *mut BdVec2) -> i32;
    pub fn bd_polygon_new(points: *const BdVec2, n: u64, is_static: i32, out_particles: *mut BdParticle,
                          out_links: *mut BdLink, out_center: *mut BdVec2) -> i32;
}

/// The library's message for the last failed call on this thread.
pub fn last_error() -> String {
    unsafe {
        let p = bd_last_error();
        if p.is_null() {
            String::new()
        } else {
            std::ffi::CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    }
}
