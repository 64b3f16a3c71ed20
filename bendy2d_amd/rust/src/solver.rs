//! `Solver` (src/solver.rs:1-189 of the crate) over libbendy2d.so.
//!
//! The scene lives on the GPU between updates.  Adders send bodies to the
//! library; `update` pushes the public fields, runs the step on the device
//! and starts copying its result to pinned host memory (readback on), so the
//! getters, which lend `&Vec` as the crate's do, only wait for that copy the
//! first time they are called after an update.

use std::sync::{Mutex, OnceLock};

use crate::circle::Circle;
use crate::ffi::{self, BdBounds, BdCircle, BdLink, BdParticle, BdPolygonIn, BdPolygonInfo, BdSolver};
use crate::link::{CircleLink, ParticleLink};
use crate::particle::{v2, vec, Particle};
use crate::polygon::Polygon;
use nalgebra::Vector2;

pub enum ColliderType {
    Particle,
    Circle,
    Polygon,
}

#[derive(Debug, Copy, Clone)]
pub struct Bounds {
    pub pos: Vector2<f32>,
    pub size: Vector2<f32>,
}

pub struct Solver {
    pub gravity: Vector2<f32>,
    pub bounds: Bounds,
    pub bounds_active: bool,
    raw: *mut BdSolver,
    // the links as added (the library keeps them too; these are lent out)
    particle_links: Vec<ParticleLink>,
    circle_links: Vec<CircleLink>,
    // bodies as of the last update, filled on first use after it
    particles: OnceLock<Vec<Particle>>,
    circles: OnceLock<Vec<Circle>>,
    polygons: OnceLock<Vec<Polygon>>,
    // one call into the handle at a time from the &self getters
    ffi_lock: Mutex<()>,
}

// The handle is used by one thread at a time: &mut self methods are
// exclusive, and the &self getters serialise on ffi_lock.
unsafe impl Send for Solver {}
unsafe impl Sync for Solver {}

fn check(rc: i64) -> i64 {
    if rc < 0 {
        panic!("{}", ffi::last_error());
    }
    rc
}

impl Solver {
    /// gravity (0, 98.2), bounds (0, 0) + (100, 100), bounds_active true, one
    /// substep: the library's defaults are the crate's.
    pub fn new() -> Self {
        let raw = unsafe { ffi::bd_solver_new() };
        assert!(!raw.is_null(), "{}", ffi::last_error());
        unsafe { ffi::bd_ext_set_readback(raw, 1) };
        Solver {
            gravity: Vector2::new(0.0, 98.2),
            bounds: Bounds { pos: Vector2::new(0.0, 0.0), size: Vector2::new(100.0, 100.0) },
            bounds_active: true,
            raw,
            particle_links: Vec::new(),
            circle_links: Vec::new(),
            particles: OnceLock::new(),
            circles: OnceLock::new(),
            polygons: OnceLock::new(),
            ffi_lock: Mutex::new(()),
        }
    }

    fn changed(&mut self) {
        self.particles = OnceLock::new();
        self.circles = OnceLock::new();
        self.polygons = OnceLock::new();
    }

    pub fn add_particle(&mut self, pos: Vector2<f32>) {
        check(unsafe { ffi::bd_add_particle(self.raw, v2(pos)) });
        self.changed();
    }

    pub fn add_circle(&mut self, circle: Circle) {
        let c: BdCircle = circle.to_ffi();
        check(unsafe { ffi::bd_add_circle(self.raw, &c) });
        self.changed();
    }

    pub fn add_polygon(&mut self, polygon: Polygon) {
        let pts: Vec<BdParticle> = polygon.particles.iter().map(|p| p.to_ffi()).collect();
        let links: Vec<BdLink> = polygon.particle_links.iter().map(|l| l.link.to_ffi()).collect();
        let desc = BdPolygonIn {
            particles: pts.as_ptr(),
            n_particles: pts.len() as u64,
            particle_links: links.as_ptr(),
            n_links: links.len() as u64,
            is_static: polygon.is_static as i32,
            center: v2(polygon.center),
            scale: polygon.scale,
        };
        check(unsafe { ffi::bd_add_polygon(self.raw, &desc) });
        self.changed();
    }

    pub fn add_particle_link(&mut self, link: ParticleLink) {
        let l = link.link.to_ffi();
        check(unsafe { ffi::bd_add_particle_link(self.raw, &l) });
        self.particle_links.push(link);
    }

    pub fn add_circle_link(&mut self, link: CircleLink) {
        let l = link.link.to_ffi();
        check(unsafe { ffi::bd_add_circle_link(self.raw, &l) });
        self.circle_links.push(link);
    }

    pub fn get_particle_len(&self) -> usize {
        unsafe { ffi::bd_get_particle_len(self.raw) as usize }
    }
    pub fn get_circles_len(&self) -> usize {
        unsafe { ffi::bd_get_circles_len(self.raw) as usize }
    }
    pub fn get_polygons_len(&self) -> usize {
        unsafe { ffi::bd_get_polygons_len(self.raw) as usize }
    }

    pub fn get_particle_links(&self) -> &Vec<ParticleLink> {
        &self.particle_links
    }
    pub fn get_circle_links(&self) -> &Vec<CircleLink> {
        &self.circle_links
    }

    pub fn get_particles(&self) -> &Vec<Particle> {
        self.particles.get_or_init(|| {
            let _g = self.ffi_lock.lock().unwrap();
            let n = self.get_particle_len();
            let mut out = vec![BdParticle::default(); n];
            unsafe { ffi::bd_get_particles(self.raw, out.as_mut_ptr(), n as u64) };
            out.iter().map(Particle::from_ffi).collect()
        })
    }

    pub fn get_circles(&self) -> &Vec<Circle> {
        self.circles.get_or_init(|| {
            let _g = self.ffi_lock.lock().unwrap();
            let n = self.get_circles_len();
            let mut out = vec![BdCircle::default(); n];
            unsafe { ffi::bd_get_circles(self.raw, out.as_mut_ptr(), n as u64) };
            out.iter().map(Circle::from_ffi).collect()
        })
    }

    pub fn get_polygons(&self) -> &Vec<Polygon> {
        self.polygons.get_or_init(|| {
            let _g = self.ffi_lock.lock().unwrap();
            (0..self.get_polygons_len() as u64)
                .map(|i| unsafe {
                    let mut info = BdPolygonInfo::default();
                    ffi::bd_get_polygon(self.raw, i, &mut info);
                    let mut pts = vec![BdParticle::default(); info.n_particles as usize];
                    let mut links = vec![BdLink::default(); info.n_links as usize];
                    ffi::bd_get_polygon_particles(self.raw, i, pts.as_mut_ptr(), info.n_particles);
                    ffi::bd_get_polygon_links(self.raw, i, links.as_mut_ptr(), info.n_links);
                    let mut p = Polygon::assemble(&pts, &links, info.is_static != 0, vec(info.center));
                    p.scale = info.scale;
                    p
                })
                .collect()
        })
    }

    pub fn get_particle(&self, index: usize) -> Option<&Particle> {
        self.get_particles().get(index)
    }
    pub fn get_circle(&self, index: usize) -> Option<&Circle> {
        self.get_circles().get(index)
    }
    pub fn get_polygon(&self, index: usize) -> Option<&Polygon> {
        self.get_polygons().get(index)
    }

    /// The crate's step (gravity, links, circle and polygon collisions,
    /// bounds, Verlet) on the GPU.  A link with particle_a >= particle_b or an
    /// index out of range panics, as the crate's does, but before anything
    /// has moved.
    pub fn update(&mut self, dt: f32) {
        unsafe {
            ffi::bd_set_gravity(self.raw, v2(self.gravity));
            ffi::bd_set_bounds(self.raw, BdBounds { pos: v2(self.bounds.pos), size: v2(self.bounds.size) });
            ffi::bd_set_bounds_active(self.raw, self.bounds_active as i32);
            check(ffi::bd_update(self.raw, dt) as i64);
        }
        self.changed();
    }
}

impl Default for Solver {
    fn default() -> Self {
        Solver::new()
    }
}

impl Clone for Solver {
    fn clone(&self) -> Self {
        let raw = {
            let _g = self.ffi_lock.lock().unwrap();
            unsafe { ffi::bd_solver_clone(self.raw) }
        };
        assert!(!raw.is_null(), "{}", ffi::last_error());
        unsafe { ffi::bd_ext_set_readback(raw, 1) };
        Solver {
            gravity: self.gravity,
            bounds: self.bounds,
            bounds_active: self.bounds_active,
            raw,
            particle_links: self.particle_links.clone(),
            circle_links: self.circle_links.clone(),
            particles: OnceLock::new(),
            circles: OnceLock::new(),
            polygons: OnceLock::new(),
            ffi_lock: Mutex::new(()),
        }
    }
}

impl std::fmt::Debug for Solver {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        f.debug_struct("Solver")
            .field("gravity", &self.gravity)
            .field("bounds", &self.bounds)
            .field("bounds_active", &self.bounds_active)
            .field("particles", self.get_particles())
            .field("particle_links", &self.particle_links)
            .field("circles", self.get_circles())
            .field("circle_links", &self.circle_links)
            .field("polygons", self.get_polygons())
            .finish()
    }
}

impl Drop for Solver {
    fn drop(&mut self) {
        unsafe { ffi::bd_solver_free(self.raw) };
    }
}
