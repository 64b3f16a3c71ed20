//! A pinned-free cloth over the crate's own API: the same program builds
//! against Murrent/bendy2d and against this crate; here Solver::update runs on
//! the GPU.
use bendy2d::circle::Circle;
use bendy2d::link::{Link, ParticleLink};
use bendy2d::particle::Particle;
use bendy2d::polygon::Polygon;
use bendy2d::solver::Solver;
use nalgebra::Vector2;

fn main() {
    let mut solver = Solver::new();
    let n = 16;
    for r in 0..n {
        for c in 0..n {
            solver.add_particle(Vector2::new(20.0 + c as f32, 10.0 + r as f32));
        }
    }
    for r in 0..n {
        for c in 0..n {
            let i = r * n + c;
            if c + 1 < n {
                solver.add_particle_link(ParticleLink { link: Link { particle_a: i, particle_b: i + 1, target_distance: 1.0 } });
            }
            if r + 1 < n {
                solver.add_particle_link(ParticleLink { link: Link { particle_a: i, particle_b: i + n, target_distance: 1.0 } });
            }
        }
    }
    solver.add_circle(Circle { point: Particle::new(Vector2::new(50.0, 40.0)), radius: 3.0 });
    solver.add_polygon(Polygon::circle(6.0, Vector2::new(70.0, 30.0), 8, false));
    for _frame in 0..600 {
        solver.update(1.0 / 60.0);
        let lowest = solver.get_particles().iter().map(|p| p.pos.y).fold(f32::MIN, f32::max);
        let _ = lowest;
    }
    println!("{} particles, {} circles, {} polygons", solver.get_particle_len(), solver.get_circles_len(),
             solver.get_polygons_len());
}
