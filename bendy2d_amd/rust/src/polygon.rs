//! `Polygon` (src/polygon.rs:1-238 of the crate): a soft ring of particles
//! held by internal links.

use crate::common::line_intersection;
use crate::ffi::{self, BdLink, BdParticle, BdVec2};
use crate::link::{Link, ParticleLink};
use crate::particle::{v2, vec, Particle};
use crate::solver::Bounds;
use nalgebra::Vector2;

#[derive(Debug, Clone)]
pub struct Polygon {
    pub particles: Vec<Particle>,
    pub particle_links: Vec<ParticleLink>,
    pub is_static: bool,
    pub center: Vector2<f32>,
    pub scale: f32,
}

impl Polygon {
    /// `point_count` points on a circle, each linked across the ring at a
    /// third and two thirds of the way round.  Built by the library's host
    /// constructor (bd_polygon_circle), which is bit-identical to the crate's.
    pub fn circle(radius: f32, pos: Vector2<f32>, point_count: usize, is_static: bool) -> Self {
        let mut pts = vec![BdParticle::default(); point_count];
        let mut links = vec![BdLink::default(); 2 * point_count];
        let mut center = BdVec2::default();
        let rc = unsafe {
            ffi::bd_polygon_circle(radius, v2(pos), point_count as u64, is_static as i32, pts.as_mut_ptr(),
                                   links.as_mut_ptr(), &mut center)
        };
        assert!(rc == 0, "{}", ffi::last_error());
        Self::assemble(&pts, &links, is_static, vec(center))
    }

    /// A ring through `points`, consecutive points linked (bd_polygon_new).
    pub fn new(points: Vec<Vector2<f32>>, is_static: bool) -> Self {
        let input: Vec<BdVec2> = points.iter().map(|p| v2(*p)).collect();
        let mut pts = vec![BdParticle::default(); points.len()];
        let mut links = vec![BdLink::default(); points.len()];
        let mut center = BdVec2::default();
        let rc = unsafe {
            ffi::bd_polygon_new(input.as_ptr(), input.len() as u64, is_static as i32, pts.as_mut_ptr(),
                                links.as_mut_ptr(), &mut center)
        };
        assert!(rc == 0, "{}", ffi::last_error());
        Self::assemble(&pts, &links, is_static, vec(center))
    }

    pub(crate) fn assemble(pts: &[BdParticle], links: &[BdLink], is_static: bool, center: Vector2<f32>) -> Self {
        Polygon {
            particles: pts.iter().map(Particle::from_ffi).collect(),
            particle_links: links.iter().map(|l| ParticleLink { link: Link::from_ffi(l) }).collect(),
            is_static,
            center,
            scale: 1.0,
        }
    }

    /// Verlet step of every point, after the centre is recomputed; a static
    /// polygon does not move.
    pub fn update(&mut self, dt: f32) {
        if self.is_static {
            return;
        }
        self.calc_center();
        self.particles.iter_mut().for_each(|p| p.update(dt));
    }

    pub fn solve_bounds(&mut self, bounds: Bounds) {
        self.particles.iter_mut().for_each(|p| p.solve_bounds(bounds));
    }

    /// Points of each polygon pushed out of the other's edges, this one's
    /// edges first.
    pub fn solve_polygon(&mut self, other: &mut Polygon) {
        self.solve_polygon_single(other);
        other.solve_polygon_single(self);
    }

    /// Each edge (i, i+1) of this polygon, as it was when the edge's turn
    /// came, against every point of `other`; the last hit on an edge decides
    /// where its ends go.
    pub fn solve_polygon_single(&mut self, other: &mut Polygon) {
        let n = self.particles.len();
        for i in 0..n {
            let j = (i + 1) % n;
            let (a, b) = (self.particles[i], self.particles[j]);
            for q in other.particles.iter_mut() {
                if let Some((new_a, new_b, new_q)) = self.resolve_line_intersection(&a, &b, q, other.center) {
                    self.particles[i].pos = new_a;
                    self.particles[j].pos = new_b;
                    q.pos = new_q;
                }
            }
        }
    }

    /// Where the edge a-b and the point q go when the segment from q to the
    /// other polygon's centre crosses the edge: the edge ends move two thirds
    /// of q's depth along the inward normal, split by their distances from
    /// the crossing, and q moves onto the edge along that normal.
    fn resolve_line_intersection(
        &mut self,
        point_a: &Particle,
        point_b: &Particle,
        others_point: &mut Particle,
        other_center: Vector2<f32>,
    ) -> Option<(Vector2<f32>, Vector2<f32>, Vector2<f32>)> {
        let hit = line_intersection((point_a.pos, point_b.pos), (others_point.pos, other_center))?;
        let along = (point_b.pos - point_a.pos).normalize();
        let centre_on_line = (along.dot(&(self.center - hit)) / along.dot(&along)) * along;
        let inward = (self.center - (hit + centre_on_line)).normalize();
        let to_a = (hit - point_a.pos).magnitude();
        let to_b = (hit - point_b.pos).magnitude();
        let length = to_a + to_b;
        let share_a = to_b / length;
        let share_b = to_a / length;
        let depth = hit - others_point.pos;
        let depth_on_normal = (inward.dot(&depth) / inward.dot(&inward)) * inward;
        let edge_move = (depth_on_normal / 3.0) * 2.0;
        let new_a = point_a.pos - share_a * edge_move;
        let new_b = point_b.pos - share_b * edge_move;
        let ray_end = others_point.pos - inward * 10000.0;
        line_intersection((point_a.pos, point_b.pos), (others_point.pos, ray_end)).map(|q| (new_a, new_b, q))
    }

    /// The centre, then every link in order.
    pub fn solve_links(&mut self) {
        self.calc_center();
        for link in self.particle_links.iter_mut() {
            link.solve(&mut self.particles);
        }
    }

    pub fn add_force_v2(&mut self, force: Vector2<f32>) {
        self.particles.iter_mut().for_each(|p| p.add_force_v2(force));
    }

    /// Mean of the points, summed in index order.
    fn calc_center(&mut self) {
        let mut sum = Vector2::new(0.0, 0.0);
        for p in &self.particles {
            sum += p.pos;
        }
        self.center = sum / self.particles.len() as f32;
    }
}
