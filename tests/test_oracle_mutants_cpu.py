"""Planted mistakes: each case copies the oracle to a temporary directory,
changes one line the way a plausible slip would (a dropped factor, a sign,
clamp before power, a shifted centre, a wrong constant), and runs the pin
that guards that line against the copy.  The pin must fail on the copy (it
passes on the real oracle in the main suite), so every oracle function
listed here is pinned by something that notices such a mistake."""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

import oracle_js as oj

pytestmark = pytest.mark.skipif(not oj.node_available(), reason="node is required for the oracle")

ROOT = Path(__file__).resolve().parent.parent
PINS = "tests/test_oracle_pins_cpu.py"
OLD = "tests/test_oracle_cpu.py"

# (oracle file, exact text, planted text, pin test that must catch it)
MUTANTS = [
    ("js/doe.js", "if (fb > 0 && a > 0) na = t + fb * (a - t);",
     "if (fb > 0 && a > 0) na = a + fb * (t - a);", f"{PINS}::test_hio_step_closed_form"),
    ("js/doe.js", "const fb = it >= pick(o.escape, 0) ? o.feedback : 0;",
     "const fb = it > pick(o.escape, 0) ? o.feedback : 0;", f"{PINS}::test_hio_escape_period"),
    ("js/doe.js", "weights[k] *= Math.pow(Math.min(Math.max(ratio, 0.25), 4), beta);",
     "weights[k] *= Math.min(Math.max(Math.pow(ratio, beta), 0.25), 4);",
     f"{PINS}::test_far_field_wgs_weights"),
    ("js/doe.js", "const norm = wsum > 0 ? Math.sqrt(tsum / wsum) : 1;",
     "const norm = wsum > 0 ? Math.sqrt(wsum / tsum) : 1;", f"{PINS}::test_far_field_wgs_weights"),
    ("js/doe.js", "weights[k] = Math.min(Math.max(weights[k], 1e-4), 1e4);",
     "weights[k] = Math.min(Math.max(weights[k], 1e-4), 2);", f"{PINS}::test_far_field_wgs_weights"),
    ("js/doe.js", "bestPhase = Float64Array.from(result.phase);",
     "bestPhase = Float64Array.from(sub.initPhase);", f"{PINS}::test_best_iterate_off_by_one"),
    ("js/doe.js", "        re[k] = t;\n        im[k] = 0;",
     "        re[k] = 0;\n        im[k] = 0;", f"{PINS}::test_far_field_zero_amplitude_branch"),
    ("js/doe.js", "out[i + j * nx] = ap[i + j * nx] * Math.exp(-(x * x + y * y) / (w * w));",
     "out[i + j * nx] = ap[i + j * nx] * Math.exp(-(x * x + y * y) / (2 * w * w));",
     f"{PINS}::test_illumination_profiles"),
    ("js/field.js", "else v = 0.5 * (1 + Math.cos((Math.PI * (r - radius + soft)) / soft));",
     "else v = 0.5 * (1 - Math.cos((Math.PI * (r - radius + soft)) / soft));",
     f"{PINS}::test_illumination_profiles"),
    ("js/propagate.js", "softWindow(Math.abs(fx), limX, limX * (1 - softEdge))",
     "softWindow(Math.abs(fx), limX, limX * softEdge)", f"{PINS}::test_band_limit_soft_window"),
    ("js/propagate.js", "return 0.5 * (1 + Math.cos((Math.PI * (v - inner)) / (lim - inner)));",
     "return 0.5 * (1 + Math.cos((Math.PI * (v - inner)) / lim));", f"{PINS}::test_band_limit_soft_window"),
    ("js/metrics.js", "const C2 = 0.03 * L * (0.03 * L);", "const C2 = 0.3 * L * (0.03 * L);",
     f"{PINS}::test_ssim_affine_windows"),
    ("js/metrics.js", "const C1 = 0.01 * L * (0.01 * L);", "const C1 = 0.1 * L * (0.01 * L);",
     f"{PINS}::test_ssim_piecewise_constant"),
    ("js/metrics.js", "const wt = o.weight === 'target' ? st / n : 1;",
     "const wt = o.weight === 'target' ? sp / n : 1;", f"{PINS}::test_ssim_piecewise_constant"),
    ("js/metrics.js", "const vp = Math.abs(spp / n - mp * mp);", "const vp = Math.abs(spp / n - mp);",
     f"{PINS}::test_ssim_affine_windows"),
    ("js/metrics.js", "return (sr * sr + si * si) / (n * p);", "return (sr * sr + si * si) / p;",
     f"{PINS}::test_zero_order_fraction"),
    ("js/metrics.js", "const x = (i - nx / 2) * dx - x0;", "const x = (i - (nx - 1) / 2) * dx - x0;",
     f"{PINS}::test_rms_radius_gaussian"),
    ("js/metrics.js", "if (v >= cut) cells[ci + cj * cols] += v;", "if (v >= 0) cells[ci + cj * cols] += v;",
     f"{PINS}::test_spot_uniformity"),
    ("js/metrics.js", "return { uniformity: mx + mn > 0 ? (2 * mn) / (mx + mn) : 1,",
     "return { uniformity: mx + mn > 0 ? mn / mx : 1,", f"{PINS}::test_spot_uniformity"),
    ("js/field.js", "const fx = ((i - nx2 / 2) * dx2) / this.dx + this.nx / 2 - 0.5;",
     "const fx = ((i - nx2 / 2) * dx2) / this.dx + this.nx / 2;",
     f"{PINS}::test_resample_bilinear_reproduces_affine_planes"),
    ("js/field.js", "    const off = (1 - s) / 2;", "    const off = 1 - s;",
     f"{PINS}::test_sample_image_closed_forms"),
    ("js/field.js", "acc += fill * (area - (hix - lox) * (hiy - loy));", "acc += fill * area;",
     f"{PINS}::test_sample_image_closed_forms"),
    # pins of the older suite (FFT, ASM, Fresnel, RNG)
    ("js/fft.js", "const ang = (-2 * Math.PI * k) / len;", "const ang = (2 * Math.PI * k) / len;",
     f"{OLD}::test_fft2d_matches_numpy"),
    ("js/propagate.js", "        const ph = 2 * Math.PI * z * Math.sqrt(arg);\n        re[k] = w * Math.cos(ph);",
     "        const ph = Math.PI * z * Math.sqrt(arg);\n        re[k] = w * Math.cos(ph);",
     f"{OLD}::test_asm_matches_shifted_numpy_pipeline"),
    ("js/propagate.js", "const amp = (dx * dy) / (lambda * az);", "const amp = (dx * dy) / (lambda * z * z);",
     f"{OLD}::test_fresnel_brute_force_integral"),
    ("js/doe.js", "    s ^= s >>> 17;", "    s ^= s >>> 15;", f"{OLD}::test_random_phase_known_answers"),
    ("js/propagate.js", "const kz = k * z - (Math.PI * (nx + ny)) / 2;", "const kz = k * z;",
     f"{OLD}::test_fresnel_brute_force_non_square"),
]


def _id(m):
    return f"{Path(m[0]).stem}:{m[3].split('::')[1]}:{m[1][:28]}"


@pytest.mark.parametrize("mutant", MUTANTS, ids=[_id(m) for m in MUTANTS])
def test_pin_catches_planted_mistake(mutant, tmp_path):
    rel, old, new, pin = mutant
    copy = tmp_path / "oracle"
    shutil.copytree(ROOT / "oracle", copy, ignore=shutil.ignore_patterns("_ref", "node_modules"))
    src = (copy / rel).read_text()
    assert src.count(old) == 1, f"the guarded line must occur exactly once in {rel}: {old!r}"
    (copy / rel).write_text(src.replace(old, new))
    env = dict(os.environ, DOE_ORACLE_DIR=str(copy))
    proc = subprocess.run([sys.executable, "-m", "pytest", pin, "-x", "-q", "-p", "no:cacheprovider"],
                          cwd=str(ROOT), env=env, capture_output=True, text=True, timeout=600)
    # exit status 1: tests ran and failed (not a collection or usage error)
    assert proc.returncode == 1, proc.stdout[-2000:] + proc.stderr[-2000:]
    assert "assert" in proc.stdout, proc.stdout[-2000:]
