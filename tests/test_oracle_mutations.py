"""Plants plausible mistakes in copies of the CPU restatement and checks that
the pinning tests catch every one of them (the oracle is only as good as
the checks that would notice it being wrong).  CPU only."""
import ctypes as C
import os
import subprocess

import pytest

import oracle_ffi as O
import test_oracle_pinning as P
import test_procgen_oracle as PG

SOURCES = ("vt_oracle.c", "vt_procgen_oracle.c")


def _gen(check):
    return lambda: check(PG.gen512_data())


# (name, original text, mutated text, pinning checks expected to fail[, file])
MUTATIONS = [
    ("anchor 0.999 -> 1.0", "<< s) * 0.999f;", "<< s) * 1.0f;",
     [P.test_trace_matches_amanatides_woo_with_drift, P.test_empty_world_steps_and_exit]),
    ("tie rule z>y>x -> x first", "m = t[0] < t[1] ? (t[0] < t[2] ? 0 : 2)",
     "m = t[0] <= t[1] ? (t[0] <= t[2] ? 0 : 2)", [P.test_exact_tie_skips_the_corner_blocks]),
    ("grid index transposed",
     "return w->grid[(uint64_t)bx + (uint64_t)w->dim * ((uint64_t)by + (uint64_t)w->dim * bz)];",
     "return w->grid[(uint64_t)bz + (uint64_t)w->dim * ((uint64_t)by + (uint64_t)w->dim * bx)];",
     [P.test_trace_matches_bruteforce_first_hit]),
    ("face sign flipped", "int face = m == 0 ? 2 - posv[0]", "int face = m == 0 ? 1 + posv[0]",
     [P.test_first_hit_face_and_position, P.test_hit_face_is_the_entered_face]),
    ("entry epsilon sign", "out[i] = (o[i] + d[i] * tc) - 0.001f;",
     "out[i] = (o[i] + d[i] * tc) + 0.001f;", [P.test_entry_origin]),
    ("pixel centre +0.5", "float u = (float)px / (float)W;", "float u = ((float)px + 0.5f) / (float)W;",
     [P.test_raygen_matches_closed_form]),
    ("aspect inverted", "v = v * ((float)H / (float)W);", "v = v * ((float)W / (float)H);",
     [P.test_raygen_matches_closed_form]),
    ("3-component normalize", "s = s + r[3] * r[3];", "s = s + 0.0f;",
     [P.test_raygen_centre_pixel_four_component_normalize]),
    ("entity t >= 0 check added", "if (tf >= tn && prevD >= tf) {",
     "if (tf >= tn && tf >= 0.0f && prevD >= tf) {", [P.test_entities_closed_form]),
    ("pool in-block index x+8z+64y",
     "(uint64_t)(((bz % 8) * 8 + (by % 8)) * 8);", "(uint64_t)(((by % 8) * 8 + (bz % 8)) * 8);",
     [P.test_brickmap_layout_renders_like_dense]),
    ("present texPos without +0.5",
     "float tp[2] = {((float)x + 0.5f) / (float)W, ((float)y + 0.5f) / (float)H};",
     "float tp[2] = {(float)x / (float)W, (float)y / (float)H};",
     [P.test_present_crosshair_known_sizes]),
    ("vignette exponent 0.6", "float g = powf(ux * uy * 15.0f, 0.6f * 0.3f);",
     "float g = powf(ux * uy * 15.0f, 0.6f);",
     [lambda: P.test_present_vignette_closed_form(333, 77)]),
    ("rgba8 byte order BGRA", "r |= (uint32_t)floorf(k * 255.0f + 0.5f) << (8 * i);",
     "r |= (uint32_t)floorf(k * 255.0f + 0.5f) << (8 * (i < 3 ? 2 - i : 3));",
     [P.test_present_rgba8_store]),
    ("simplex far-corner constant sign", "const float K2 = (float)(-2 * (1 - 2 * G2)",
     "const float K2 = (float)(2 * (1 - 2 * G2)",
     [PG.test_simplex_zero_on_lattice, PG.test_simplex_equals_vertex_sum], "vt_procgen_oracle.c"),
    ("simplex middle corners swapped", "n1 = (b * b) * (b * b) * fnl_grad(seed, ip, jp + FNL_PRIME_Y, x1, y1);",
     "n1 = (b * b) * (b * b) * fnl_grad(seed, ip + FNL_PRIME_X, jp, x1, y1);",
     [PG.test_simplex_equals_vertex_sum], "vt_procgen_oracle.c"),
    ("simplex hash shift 13", "h ^= (uint32_t)((int32_t)h >> 15);",
     "h ^= (uint32_t)((int32_t)h >> 13);", [PG.test_simplex_equals_vertex_sum],
     "vt_procgen_oracle.c"),
    ("fbm octave seed fixed", "float noise = or_fnl_simplex2(seed++, x, y);",
     "float noise = or_fnl_simplex2(seed, x, y);",
     [PG.test_fbm_equals_vertex_sum_and_is_bounded], "vt_procgen_oracle.c"),
    ("fbm skew dropped", "  x += t;\n  y += t;\n  /* _fnlGenFractalFBM2D */",
     "  /* _fnlGenFractalFBM2D */", [PG.test_fbm_equals_vertex_sum_and_is_bounded],
     "vt_procgen_oracle.c"),
    ("procgen closing draw dropped", "      (void)lcg(&s);\n    }\n  }",
     "    }\n  }", [_gen(PG.test_lcg_replay_first_rows)], "vt_procgen_oracle.c"),
    ("procgen top layer from the first draw", "else if (h == vh - 1 && h > 15)",
     "else if (h == vh && h > 15)", [_gen(PG.test_column_layers)], "vt_procgen_oracle.c"),
    ("procgen chunk allocation y-major",
     "for (uint64_t x = 0; x < dim; x++)\n    for (uint64_t z = 0; z < dim; z++)\n      for (uint64_t y = 0; y < 16; y++)",
     "for (uint64_t y = 0; y < 16; y++)\n    for (uint64_t x = 0; x < dim; x++)\n      for (uint64_t z = 0; z < dim; z++)",
     [_gen(PG.test_base_layer_allocation_order)], "vt_procgen_oracle.c"),
]


@pytest.mark.parametrize("mutation", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_pinning_catches_planted_mistake(tmp_path, mutation):
    name, orig, mut, checks = mutation[:4]
    target = mutation[4] if len(mutation) > 4 else "vt_oracle.c"
    for f in SOURCES + ("vt_oracle.h",):
        text = open(os.path.join(O.ORACLE_DIR, f)).read()
        if f == target:
            assert text.count(orig) == 1, f"mutation anchor not found once: {orig}"
            text = text.replace(orig, mut)
        (tmp_path / f).write_text(text)
    so = tmp_path / "liboracle.so"
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-o",
                    str(so), *[str(tmp_path / f) for f in SOURCES], "-lm", "-lpthread"],
                   check=True)
    saved_dir, saved_lib = O.ORACLE_DIR, O._L
    try:
        O.ORACLE_DIR, O._L = str(tmp_path), None
        caught = False
        for check in checks:
            try:
                if "c0_render" in check.__code__.co_varnames[:check.__code__.co_argcount]:
                    check(P.c0_render.__wrapped__())
                else:
                    check()
            except AssertionError:
                caught = True
                break
        assert caught, f"planted mistake '{name}' passed every pinning check"
    finally:
        O.ORACLE_DIR, O._L = saved_dir, saved_lib
