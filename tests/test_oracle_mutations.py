"""Plants plausible slips in copies of the oracle and checks that the oracle's
tests catch each one.

Every entry below changes one expression of oracle/oracle_impl.h the way a
restatement of the Rust source could plausibly get it wrong (a reassociation, a
reciprocal for a division, a sign, an index, a comparison, a phase or loop
order).  The copy is built with the oracle's own flags into a temporary
directory, the known-answer tests (test_oracle_kat.py) and the model tests
(test_oracle_model.py) are pointed at it, and at least one of them must fail.
The real oracle is restored afterwards.  This runs on the CPU only: the slips
are planted in the oracle, never in GPU code.
"""
import inspect
import os
import shutil
import subprocess

import pytest

import helpers  # noqa: F401  (puts oracle/ on sys.path)
import oracle
import test_oracle_kat
import test_oracle_model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

# (name, reference file:line the slip breaks, text in oracle_impl.h, replacement)
MUTATIONS = [
    ("normalize as a multiply by the reciprocal", "nalgebra normalize / src/link.rs:84",
     "static inline T(v2) F(vnormalize)(T(v2) a) { return F(vdiv)(a, F(vmag)(a)); }",
     "static inline T(v2) F(vnormalize)(T(v2) a) { return F(vmul)(a, (REAL)1.0f / F(vmag)(a)); }"),
    ("Particle bounds x-high reassociated", "src/particle.rs:34",
     "p->prev.x = (bp.x + bs.x) - v;", "p->prev.x = bp.x + (bs.x - v);"),
    ("Particle bounds y-high wrong sign", "src/particle.rs:43",
     "p->prev.y = (bp.y + bs.y) - v;", "p->prev.y = (bp.y + bs.y) + v;"),
    ("Particle bounds y-low wrong sign", "src/particle.rs:39",
     "p->prev.y = bp.y - v;", "p->prev.y = bp.y + v;"),
    ("Circle bounds x-low reassociated", "src/circle.rs:123",
     "p->prev.x = (bp.x + r) - v;", "p->prev.x = bp.x + (r - v);"),
    ("Circle bounds y-high radius added", "src/circle.rs:134-137",
     "} else if (p->pos.y > (bp.y + bs.y) - r) {", "} else if (p->pos.y > (bp.y + bs.y) + r) {"),
    ("Polygon::circle first link to i+2n/3+1", "src/polygon.rs:35",
     "(i + 2 * n / 3) % n", "(i + 2 * n / 3 + 1) % n"),
    ("Polygon::circle second link to i+n/3+1", "src/polygon.rs:55",
     "(i + n / 3) % n;", "(i + n / 3 + 1) % n;"),
    ("Polygon::new links unswapped", "src/polygon.rs:97-103",
     "        int64_t a = i, b = (i + 1) % n;\n        if (a > b) { int64_t t = a; a = b; b = t; }",
     "        int64_t a = i, b = (i + 1) % n;"),
    ("ParticleLink half applied to e first", "src/link.rs:85-86",
     "T(v2) c = F(vmul)(F(vmul)(n, e), (REAL)0.5f);", "T(v2) c = F(vmul)(n, e * (REAL)0.5f);"),
    ("CircleLink weight grouped as e*(scale*rb2)", "src/link.rs:106",
     "F(vmul)(F(vmul)(F(vmul)(n, e), scale), rb2)", "F(vmul)(n, e * (scale * rb2))"),
    ("CircleLink weights swapped", "src/link.rs:106-107",
     "ca->pt.pos = F(vsub)(ca->pt.pos, F(vmul)(F(vmul)(F(vmul)(n, e), scale), rb2));",
     "ca->pt.pos = F(vsub)(ca->pt.pos, F(vmul)(F(vmul)(F(vmul)(n, e), scale), ra2));"),
    ("Verlet pos + (vel + acc*dt*dt)", "src/particle.rs:23",
     "p->pos = F(vadd)(F(vadd)(p->pos, vel), F(vmul)(F(vmul)(p->acc, dt), dt));",
     "p->pos = F(vadd)(p->pos, F(vadd)(vel, F(vmul)(F(vmul)(p->acc, dt), dt)));"),
    ("Verlet acc*(dt*dt)", "src/particle.rs:23",
     "F(vmul)(F(vmul)(p->acc, dt), dt)", "F(vmul)(p->acc, dt * dt)"),
    ("circle pair touching counts as contact", "src/circle.rs:145",
     "    REAL rs = ci->r + cj->r;\n    if (q < rs * rs) {", "    REAL rs = ci->r + cj->r;\n    if (q <= rs * rs) {"),
    ("circle pair weights swapped", "src/circle.rs:151",
     "T(v2) di = F(vmul)(F(vmul)(F(vmul)(n, scale), ov), rj2);",
     "T(v2) di = F(vmul)(F(vmul)(F(vmul)(n, scale), ov), ri2);"),
    ("circle pair product grouped as n*(scale*ov)", "src/circle.rs:151",
     "T(v2) di = F(vmul)(F(vmul)(F(vmul)(n, scale), ov), rj2);",
     "T(v2) di = F(vmul)(F(vmul)(n, scale * ov), rj2);"),
    ("circle pairs in reverse order", "src/solver.rs:169-176",
     "for (int64_t j = i + 1; j < s->nc; j++)\n            F(circle_pair)",
     "for (int64_t j = s->nc - 1; j > i; j--)\n            F(circle_pair)"),
    ("line_intersection s > 0 (endpoint excluded)", "src/common.rs:173",
     "if (s >= 0 && s <= 1 && t >= 0 && t <= 1) {", "if (s > 0 && s <= 1 && t >= 0 && t <= 1) {"),
    ("line_intersection t < 1 (endpoint excluded)", "src/common.rs:173",
     "if (s >= 0 && s <= 1 && t >= 0 && t <= 1) {", "if (s >= 0 && s <= 1 && t >= 0 && t < 1) {"),
    ("line_intersection denominator sign", "src/common.rs:170-171",
     "REAL den = (-s2x) * s1y + s1x * s2y;", "REAL den = s2x * s1y - s1x * s2y;"),
    ("resolve_line displacement /3 as *(1/3)", "src/polygon.rs:196",
     "T(v2) third = F(vdiv)(ion, (REAL)3.0f);", "T(v2) third = F(vmul)(ion, (REAL)1.0f / (REAL)3.0f);"),
    ("resolve_line influences swapped", "src/polygon.rs:188-189",
     "REAL ia = db / dab, ib = da / dab;", "REAL ia = da / dab, ib = db / dab;"),
    ("resolve_line projection without the |n|^2 division", "src/polygon.rs:177-178",
     "F(vdot)(nl, F(vsub)(self_center, inter)) / F(vdot)(nl, nl))", "F(vdot)(nl, F(vsub)(self_center, inter)))"),
    ("polygon edge copies re-read live", "src/polygon.rs:149-151",
     "if (F(resolve_line)(self->center, pa, pb, other->pts[k].pos",
     "if (F(resolve_line)(self->center, self->pts[i].pos, self->pts[bid].pos, other->pts[k].pos"),
    ("solve_polygon one way only", "src/polygon.rs:142-145",
     "            F(polygon_single)(&s->polys[j], &s->polys[i]);", ""),
    ("calc_center divides by n - 1", "src/polygon.rs:236",
     "p->center = F(vdiv)(c, (REAL)(float)p->n);", "p->center = F(vdiv)(c, (REAL)(float)(p->n - 1));"),
    ("static polygons integrated", "src/polygon.rs:126-128",
     "        if (pg->is_static) continue;\n", ""),
    ("bounds before collisions", "src/solver.rs:112-113",
     "        F(dynamic_collisions)(s);\n        F(boundary)(s);", "        F(boundary)(s);\n        F(dynamic_collisions)(s);"),
    ("substep delta as dt / n", "src/solver.rs:107-108",
     "REAL delta = (REAL)dt * s->sub_steps_multiplier;", "REAL delta = (REAL)dt / (REAL)(float)s->sub_steps;"),
    ("default gravity sign", "src/solver.rs:36",
     "s->gravity.y = (REAL)98.2f;", "s->gravity.y = (REAL)-98.2f;"),
]


def _tests():
    """Every test function of the two suites, with its parametrisations expanded."""
    out = []
    for mod in (test_oracle_kat, test_oracle_model):
        for name, fn in inspect.getmembers(mod, inspect.isfunction):
            if not name.startswith("test_") or fn.__module__ != mod.__name__:
                continue
            marks = getattr(fn, "pytestmark", [])
            params = [m for m in marks if m.name == "parametrize"]
            if not params:
                out.append((f"{mod.__name__}.{name}", fn, {}))
                continue
            argnames, values = params[0].args[0], params[0].args[1]
            names = [a.strip() for a in argnames.split(",")] if isinstance(argnames, str) else list(argnames)
            for v in values:
                v = v if len(names) > 1 else (v,)
                out.append((f"{mod.__name__}.{name}{list(v)}", fn, dict(zip(names, v))))
    return out


def _run(fn, kw):
    if "_quiet_fp" in inspect.signature(fn).parameters:
        kw = dict(kw, _quiet_fp=None)
    import numpy as np
    with np.errstate(all="ignore"):
        fn(**kw)


def caught_by(lib_path, stop_at_first=True):
    """Names of the tests that fail against the oracle built at lib_path."""
    oracle.use_library(lib_path)
    failed = []
    try:
        for name, fn, kw in _tests():
            try:
                _run(fn, kw)
            except Exception:
                failed.append(name)
                if stop_at_first:
                    break
    finally:
        oracle.use_library(None)
    return failed


def build_mutant(src_text, out_dir):
    shutil.copy(os.path.join(ORACLE, "oracle.c"), out_dir)
    with open(os.path.join(out_dir, "oracle_impl.h"), "w") as f:
        f.write(src_text)
    lib = os.path.join(out_dir, "liboracle.so")
    cflags = ["-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-w"]
    subprocess.run(["gcc", *cflags, "-shared", "-o", lib, os.path.join(out_dir, "oracle.c"), "-lm"], check=True)
    return lib


def test_the_real_oracle_passes_every_test():
    assert caught_by(None, stop_at_first=False) == []


@pytest.mark.parametrize("name,ref,old,new", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_slip_is_caught(name, ref, old, new, tmp_path):
    src = open(os.path.join(ORACLE, "oracle_impl.h")).read()
    assert src.count(old) >= 1, f"mutation text not found: {old!r}"
    lib = build_mutant(src.replace(old, new), str(tmp_path))
    assert caught_by(lib), f"no oracle test fails when {name} ({ref})"
