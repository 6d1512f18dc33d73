"""Mutation check of the oracle pins: each typical mistake, planted in a copy
of the CPU oracle (oracle/lbm_oracle.cpp) on the CPU, must make the pin that
guards it fail.

For every mutant the source is copied to a temporary directory, one exact
substring is replaced (it must occur the stated number of times, so that an
edit of the oracle cannot turn a mutant into a silent no-op), the copy is
built as a shared library, and the named pins of tests/test_oracle_pins.py
run against it in a child process (ORACLE_LIB points the checker at the
build).  The unmodified source, built the same way, must pass the same pins.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "oracle" / "lbm_oracle.cpp"

# (id, old, new, occurrences of old, which occurrence to replace, pins that must fail)
MUTANTS = [
    ("bc_inflow_partner_at_distance_1",
     "if (lab == GEO_IN_L) stream(src, kl.f, x, xp, xp + 1, ym, y, yp, zm, z, zp);",
     "if (lab == GEO_IN_L) stream(src, kl.f, xm, x, xp, ym, y, yp, zm, z, zp);",
     1, 0, "partner_reads"),
    ("bc_outflow_partner_at_distance_1",
     "if (lab == GEO_OUT_F) stream(src, k.f, xm, x, xp, y, yp, yp + 1, zm, z, zp);",
     "if (lab == GEO_OUT_F) stream(src, k.f, xm, x, xp, ym, y, yp, zm, z, zp);",
     1, 0, "partner_reads"),
    ("bc_outflow_partner_two_planes_in",
     "else if (lab == GEO_OUT_B) stream(src, k.f, xm, x, xp, ym - 1, ym, y, zm, z, zp);",
     "else if (lab == GEO_OUT_B) stream(src, k.f, xm, x, xp, ym - 2, ym - 1, ym, zm, z, zp);",
     1, 0, "partner_reads"),
    ("bc_x_outflow_partner_at_distance_1",
     "else if (lab == 7) stream(src, k.f, xm - 1, xm, x, ym, y, yp, zm, z, zp);",
     "else if (lab == 7) stream(src, k.f, xm, x, xp, ym, y, yp, zm, z, zp);",
     1, 0, "partner_reads"),
    ("periodic_instead_of_clamped_neighbours",
     "const int xp = std::min(x + 1, X - 1), xm = std::max(x - 1, 0);",
     "const int xp = (x + 1) % X, xm = (x + X - 1) % X;",
     2, 0, "partner_reads"),
    ("backward_chimera_sign",
     "T fp = (k0 * (u * u + u) + k1 * (T(2) * u + T(1)) + k2) * T(0.5);",
     "T fp = (k0 * (u * u + u) - k1 * (T(2) * u + T(1)) + k2) * T(0.5);",
     1, 0, "cumulant"),
    ("cumulant_force_sign",
     "n[K(1, 0, 0)] = -k_(1, 0, 0);", "n[K(1, 0, 0)] = k_(1, 0, 0);",
     1, 0, "cumulant_against_gaussian_closure"),
    ("smagorinsky_constant",
     "const T C = T(0.0342);", "const T C = T(0.0360);",
     1, 0, "mrt_les"),
    ("kbc_gamma_sign",
     "T(1) / beta - (T(2) - T(1) / beta) * num / den",
     "T(1) / beta + (T(2) - T(1) / beta) * num / den",
     1, 0, "kbc_n4"),
    ("kbc_dropped_third_order_term",
     "v = D + Tr / T(6) + Qv;", "v = D + Tr / T(6);",
     1, 0, "kbc_n4"),
    ("neq_shear_rate_without_sqrt2",
     "const double G = 3.0 * std::sqrt(2.0) * std::sqrt(s2) / (2.0 * rho);",
     "const double G = 3.0 * std::sqrt(s2) / (2.0 * rho);",
     1, 0, "neq_viscosity_recovers_shear_rate"),
    ("bounce_back_not_reversed",
     "for (int q = 0; q < 27; ++q) f[26 - q] = src[DF(q, xs[CX(q) + 1], ys[CY(q) + 1], zs[CZ(q) + 1])];",
     "for (int q = 0; q < 27; ++q) f[q] = src[DF(q, xs[CX(q) + 1], ys[CY(q) + 1], zs[CZ(q) + 1])];",
     1, 0, "streaming_single_population"),
    ("statistics_dropped_rho",
     "acc[3][i] += r * r * ux * ux;", "acc[3][i] += r * ux * ux;",
     1, 0, "statistics_sums_closed_form"),
    ("pmm_quirk_on_wrong_direction",
     "out[18] *= rho;  // pmm", "out[20] *= rho;  // pmm",
     1, 0, "feq_central_moments_and_pmm_quirk"),
    ("strain_transposed_derivative",
     "s32 = T(0.5) * dy[2] + T(0.5) * dz[1];", "s32 = T(0.5) * dz[2] + T(0.5) * dy[1];",
     1, 0, "fd_strain_exact_on_linear_fields"),
]

BASE_PINS = " or ".join(sorted({m[5] for m in MUTANTS}))


def _build(name: str, text: str, out: Path) -> Path:
    d = out / name
    d.mkdir(parents=True, exist_ok=True)
    src = d / "lbm_oracle.cpp"
    src.write_text(text)
    lib = d / "liboracle.so"
    # the source includes "../include/lbm355.h": resolved from oracle/ via -I
    subprocess.run(["g++", "-O2", "-fopenmp", "-std=c++17", "-fPIC", "-shared", f"-I{ROOT / 'oracle'}",
                    "-o", str(lib), str(src)], check=True, capture_output=True)
    return lib


def _mutate(text: str, old: str, new: str, count: int, which: int) -> str:
    assert text.count(old) == count, f"mutant target occurs {text.count(old)} times, expected {count}: {old!r}"
    pos = -1
    for _ in range(which + 1):
        pos = text.index(old, pos + 1)
    return text[:pos] + new + text[pos + len(old):]


def _run_pins(lib: Path, expr: str) -> subprocess.CompletedProcess:
    env = dict(os.environ, ORACLE_LIB=str(lib), OMP_NUM_THREADS="2")
    return subprocess.run([sys.executable, "-m", "pytest", str(ROOT / "tests" / "test_oracle_pins.py"), "-q", "-x",
                           "-p", "no:cacheprovider", "-k", expr], cwd=ROOT, env=env, capture_output=True, text=True)


@pytest.fixture(scope="module")
def libs(tmp_path_factory):
    out = tmp_path_factory.mktemp("mutants")
    text = SRC.read_text()
    jobs = {"base": text}
    for name, old, new, count, which, _ in MUTANTS:
        jobs[name] = _mutate(text, old, new, count, which)
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        futs = {k: ex.submit(_build, k, v, out) for k, v in jobs.items()}
        return {k: f.result() for k, f in futs.items()}


@pytest.fixture(scope="module")
def results(libs):
    runs = {"base": BASE_PINS}
    runs.update({m[0]: m[5] for m in MUTANTS})
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        futs = {k: ex.submit(_run_pins, libs[k], expr) for k, expr in runs.items()}
        return {k: f.result() for k, f in futs.items()}


def test_unmodified_oracle_passes_the_pins(results):
    r = results["base"]
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]


@pytest.mark.parametrize("mutant", [m[0] for m in MUTANTS])
def test_planted_mistake_fails_its_pin(results, mutant):
    r = results[mutant]
    # pytest exit code 1: tests ran and at least one failed (not a usage or collection error)
    assert r.returncode == 1, f"{mutant}: pins did not fail (rc {r.returncode})\n{r.stdout[-1500:]}"
    assert "failed" in r.stdout
