//! bendy2d with its solver step on an AMD Instinct MI355X.
//!
//! The modules and public items are those of Murrent/bendy2d 0.0.1
//! (src/lib.rs:1-6), so code written against that crate builds against this
//! one unchanged.  `Solver` holds its scene on the GPU through the C ABI of
//! libbendy2d.so (include/bendy2d.h); `Solver::update` runs the reference's
//! substep (gravity, links, circle and polygon collisions, bounds, Verlet) in
//! HIP kernels with f32 results identical to the crate's.  The body types are
//! the crate's plain values; their methods keep the crate's arithmetic for
//! code that calls them directly.

pub mod circle;
pub mod common;
pub mod link;
pub mod particle;
pub mod polygon;
pub mod solver;

mod ffi;
