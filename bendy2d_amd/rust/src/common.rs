//! Geometry helpers (src/common.rs:1-52 of the crate).

use nalgebra::Vector2;

/// Intersection of segments line1 = (p1, p2) and line2 = (p3, p4), both
/// parameters in [0, 1] inclusive; the point is p1 + t (p2 - p1).
#[inline]
pub fn line_intersection(
    line1: (Vector2<f32>, Vector2<f32>),
    line2: (Vector2<f32>, Vector2<f32>),
) -> Option<Vector2<f32>> {
    let (p1, p2) = line1;
    let (p3, p4) = line2;
    let (s1x, s1y) = (p2.x - p1.x, p2.y - p1.y);
    let (s2x, s2y) = (p4.x - p3.x, p4.y - p3.y);
    let (ox, oy) = (p1.x - p3.x, p1.y - p3.y);
    let s = (-s1y * ox + s1x * oy) / (-s2x * s1y + s1x * s2y);
    let t = (s2x * oy - s2y * ox) / (-s2x * s1y + s1x * s2y);
    if (0.0..=1.0).contains(&s) && (0.0..=1.0).contains(&t) {
        Some(Vector2::new(p1.x + (t * s1x), p1.y + (t * s1y)))
    } else {
        None
    }
}

/// a projected on b.
#[inline]
pub fn proj_a_on_b(a: Vector2<f32>, b: Vector2<f32>) -> Vector2<f32> {
    a.dot(&b) / b.magnitude_squared() * b
}

/// a projected on b, clamped to the segment from 0 to b (componentwise
/// comparisons, as nalgebra's PartialOrd on vectors).
#[inline]
pub fn proj_a_on_b_clamped(a: Vector2<f32>, b: Vector2<f32>) -> Vector2<f32> {
    let proj = proj_a_on_b(a, b);
    if proj.xy() > b.xy() {
        b
    } else if proj.xy() < Vector2::zeros() {
        Vector2::zeros()
    } else {
        proj
    }
}

/// point projected on the segment `line`, clamped to it.
#[inline]
pub fn proj_point_on_line(point: Vector2<f32>, line: (Vector2<f32>, Vector2<f32>)) -> Vector2<f32> {
    let (start, end) = line;
    start + proj_a_on_b_clamped(point - start, end - start)
}
