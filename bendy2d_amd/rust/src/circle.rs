//! `Circle` (src/circle.rs:1-46 of the crate): a particle with a radius.

use crate::ffi::BdCircle;
use crate::particle::Particle;
use crate::solver::Bounds;

#[derive(Debug, Copy, Clone)]
pub struct Circle {
    pub point: Particle,
    pub radius: f32,
}

impl Circle {
    /// Reflect off the box shrunk by the radius (low side first per axis).
    pub fn solve_bounds(&mut self, bounds: Bounds) {
        let r = self.radius;
        let p = &mut self.point;
        for axis in 0..2 {
            let lo = bounds.pos[axis] + r;
            let hi = bounds.pos[axis] + bounds.size[axis] - r;
            if p.pos[axis] < lo {
                let v = p.prev_pos[axis] - p.pos[axis];
                p.prev_pos[axis] = lo - v;
                p.pos[axis] = lo;
            } else if p.pos[axis] > hi {
                let v = p.prev_pos[axis] - p.pos[axis];
                p.prev_pos[axis] = hi - v;
                p.pos[axis] = hi;
            }
        }
    }

    /// Push two overlapping circles apart, each by the other's share of the
    /// squared radii (inside a `Solver` every pair i < j runs on the GPU in
    /// the crate's order).
    pub fn solve_circle(&mut self, circle: &mut Circle) {
        let d = self.point.pos - circle.point.pos;
        let q = d.norm_squared();
        let reach = self.radius + circle.radius;
        if q < reach * reach {
            let n = d.normalize();
            let overlap = reach - q.sqrt();
            let w_self = self.radius * self.radius;
            let w_other = circle.radius * circle.radius;
            let scale = 1.0 / (w_self + w_other);
            self.point.pos += n * scale * overlap * w_other;
            circle.point.pos -= n * scale * overlap * w_self;
        }
    }

    pub(crate) fn to_ffi(&self) -> BdCircle {
        BdCircle { point: self.point.to_ffi(), radius: self.radius }
    }

    pub(crate) fn from_ffi(c: &BdCircle) -> Self {
        Circle { point: Particle::from_ffi(&c.point), radius: c.radius }
    }
}
