//! Distance constraints (src/link.rs:1-49 of the crate).

use crate::circle::Circle;
use crate::ffi::BdLink;
use crate::particle::Particle;
use nalgebra::Vector2;

#[derive(Debug, Copy, Clone)]
pub struct Link {
    // particle_a needs to be lower than particle_b (a Solver checks it before
    // anything moves and panics otherwise, as the crate's solve would)
    pub particle_a: usize,
    pub particle_b: usize,
    pub target_distance: f32,
}

#[derive(Debug, Copy, Clone)]
pub struct ParticleLink {
    pub link: Link,
}

#[derive(Debug, Copy, Clone)]
pub struct CircleLink {
    pub link: Link,
}

impl Link {
    pub(crate) fn to_ffi(&self) -> BdLink {
        BdLink { particle_a: self.particle_a as u64, particle_b: self.particle_b as u64,
                 target_distance: self.target_distance }
    }

    pub(crate) fn from_ffi(l: &BdLink) -> Self {
        Link { particle_a: l.particle_a as usize, particle_b: l.particle_b as usize,
               target_distance: l.target_distance }
    }
}

/// Both ends of a link, mutably (a < b, as the crate requires).
fn ends<T>(items: &mut [T], link: &Link) -> (&mut T, &mut T) {
    let (low, high) = items.split_at_mut(link.particle_b);
    (&mut low[link.particle_a], &mut high[0])
}

impl ParticleLink {
    /// Move both ends half the length error each along the link (inside a
    /// `Solver` this runs on the GPU, in dependency layers that keep the
    /// sequential order).
    pub fn solve(&mut self, particles: &mut Vec<Particle>) {
        let (a, b) = ends(particles, &self.link);
        let d: Vector2<f32> = a.pos - b.pos;
        let length = d.magnitude();
        let n = d.normalize();
        let error = length - self.link.target_distance;
        a.pos -= n * error * 0.5;
        b.pos += n * error * 0.5;
    }
}

impl CircleLink {
    /// As ParticleLink::solve, each end moving by the other's share of the
    /// squared radii.
    pub fn solve(&mut self, circles: &mut Vec<Circle>) {
        let (a, b) = ends(circles, &self.link);
        let d: Vector2<f32> = a.point.pos - b.point.pos;
        let length = d.magnitude();
        let n = d.normalize();
        let wa = a.radius * a.radius;
        let wb = b.radius * b.radius;
        let scale = 1.0 / (wa + wb);
        let error = length - self.link.target_distance;
        a.point.pos -= n * error * scale * wb;
        b.point.pos += n * error * scale * wa;
    }
}
