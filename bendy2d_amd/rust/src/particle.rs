//! `Particle` (src/particle.rs:4-60 of the crate): a point with Verlet state.

use crate::ffi::{BdParticle, BdVec2};
use crate::solver::Bounds;
use nalgebra::Vector2;

#[derive(Debug, Copy, Clone)]
pub struct Particle {
    pub pos: Vector2<f32>,
    pub prev_pos: Vector2<f32>,
    pub(crate) acc: Vector2<f32>,
}

impl Particle {
    /// At rest at `pos` with no accumulated force.
    pub fn new(pos: Vector2<f32>) -> Self {
        Particle { pos, prev_pos: pos, acc: Vector2::zeros() }
    }

    /// One Verlet step of `dt` (inside a `Solver` this runs on the GPU):
    /// the new position is (pos + velocity) + (acc * dt) * dt, in that order.
    pub fn update(&mut self, dt: f32) {
        let velocity = self.pos - self.prev_pos;
        let step = self.acc * dt * dt;
        self.prev_pos = self.pos;
        self.pos = self.pos + velocity + step;
        self.acc = Vector2::zeros();
    }

    /// Reflect off the box `bounds`: per axis the low side is tested first and
    /// the high side only when the low one did not fire; the position goes to
    /// the side and the previous position is mirrored through it.
    pub fn solve_bounds(&mut self, bounds: Bounds) {
        let lo = bounds.pos;
        for axis in 0..2 {
            let hi = lo[axis] + bounds.size[axis];
            if self.pos[axis] < lo[axis] {
                let v = self.prev_pos[axis] - self.pos[axis];
                self.prev_pos[axis] = lo[axis] - v;
                self.pos[axis] = lo[axis];
            } else if self.pos[axis] > hi {
                let v = self.prev_pos[axis] - self.pos[axis];
                self.prev_pos[axis] = hi - v;
                self.pos[axis] = hi;
            }
        }
    }

    pub fn add_force(&mut self, force_x: f32, force_y: f32) {
        self.acc += Vector2::new(force_x, force_y);
    }

    pub fn add_force_v2(&mut self, force: Vector2<f32>) {
        self.acc += force;
    }

    pub fn add_force_towards(&mut self, point: Vector2<f32>, force: f32) {
        let direction: Vector2<f32> = (point - self.pos).normalize();
        self.acc += direction * force;
    }

    pub(crate) fn to_ffi(&self) -> BdParticle {
        BdParticle { pos: v2(self.pos), prev_pos: v2(self.prev_pos), acc: v2(self.acc) }
    }

    pub(crate) fn from_ffi(p: &BdParticle) -> Self {
        Particle { pos: vec(p.pos), prev_pos: vec(p.prev_pos), acc: vec(p.acc) }
    }
}

pub(crate) fn v2(v: Vector2<f32>) -> BdVec2 {
    BdVec2 { x: v.x, y: v.y }
}

pub(crate) fn vec(v: BdVec2) -> Vector2<f32> {
    Vector2::new(v.x, v.y)
}
