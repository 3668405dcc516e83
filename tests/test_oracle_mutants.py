"""Sensitivity of the oracle's pins: copies of oracle/oracle.cpp with one
plausible transcription mistake each (planted in the forward and, where it has
one, in the matching adjoint line, so the adjoint's transposition tests stay
green) are built and the mathematical pins are run against them.  Every mutant
must fail at least one pin.

The mutants cover the fluid-solid RHS terms that are easy to mistype: cross
metric terms of the four kernels, stiffness indices, PML memory signs and
projections, the conservative flux projections, the ghost flux, the interface
rotation and the filter's averaging (include/pml_fs_kernels.cuh:333-827,
src/pml_fs_kernels.cu:385-525).
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "oracle.cpp")
PINS = [os.path.join(ROOT, "tests", f) for f in ("test_oracle_fs_pins.py", "test_oracle_pins.py")]

# name -> [(text, replacement, n)]: the first n occurrences are replaced (the
# fluid-solid functions precede the solid-only ones in oracle.cpp)
MUTANTS = {
    "unmutated": [],
    # velocity_fluid: xi_z for ze_x in the x rate (forward and adjoint)
    "fluid_v_cross_metric": [("rvx = -(xx * dpx + zx * dpz);", "rvx = -(xx * dpx + xz * dpz);", 2),
                             ("const double dpzb = -zx * rvxb", "const double dpzb = -xz * rvxb", 1)],
    # stress_solid (fluid-solid grid): C11 and C13 swapped in the sigma_xx rate
    "solid_sxx_c11_c13": [("double ksx = c11 * (xx * dvxx + zx * dvxz) + c13 * (xz * dvzx + zz * dvzz);\n"
                           "            double ksz = c13 * (xx * dvxx + zx * dvxz) + c33 * (xz * dvzx + zz * dvzz);\n"
                           "            double ksxz;",
                           "double ksx = c13 * (xx * dvxx + zx * dvxz) + c11 * (xz * dvzx + zz * dvzz);\n"
                           "            double ksz = c13 * (xx * dvxx + zx * dvxz) + c33 * (xz * dvzx + zz * dvzz);\n"
                           "            double ksxz;", 1),
                          ("const double exxb = c11 * qsx + c13 * qsz, ezzb = c13 * qsx + c33 * qsz;",
                           "const double exxb = c13 * qsx + c13 * qsz, ezzb = c11 * qsx + c33 * qsz;", 1)],
    # velocity_fluid: sign of the xi memory term in the velocity rate
    "fluid_v_memory_sign": [("Kd[FVX][idx] += ((invb - 1.0) * Dx - invb * S[FXVX][idx]) / rho;",
                             "Kd[FVX][idx] += ((invb - 1.0) * Dx + invb * S[FXVX][idx]) / rho;", 1),
                            ("Sb[FXVX][idx] += -invb / rho * qvx - apdp * qVx;",
                             "Sb[FXVX][idx] += +invb / rho * qvx - apdp * qVx;", 1),
                            ("kvx += (invb - 1.0) * Dx - invb * S[FXVX][idx];",
                             "kvx += (invb - 1.0) * Dx + invb * S[FXVX][idx];", 1)],
    # stress_fluid: ze_x and xi_z exchanged in the divergence
    "fluid_p_cross_metric": [("Kd[FP][idx] = -Km * (xx * dvxx + zx * dvxz + xz * dvzx + zz * dvzz);",
                              "Kd[FP][idx] = -Km * (xx * dvxx + xz * dvxz + zx * dvzx + zz * dvzz);", 1)],
    # stress_solid: ze_x and xi_z exchanged in the shear rate
    "solid_sxz_cross_metric": [("ksxz = c55 * (xz * dvxx + zz * dvxz + xx * dvzx + zx * dvzz);",
                                "ksxz = c55 * (zx * dvxx + zz * dvxz + xx * dvzx + xz * dvzz);", 1)],
    # velocity_solid: ze_x and xi_z exchanged in the z rate
    "solid_v_cross_metric": [("rvz = xx * e + zx * f + xz * c + zz * d;", "rvz = xx * e + xz * f + zx * c + zz * d;", 1)],
    # velocity_solid: sign of the xi memory term
    "solid_v_memory_sign": [("kvx += ((invb - 1.0) * Dxx - invb * S[SXVX][idx]) / rho;",
                             "kvx += ((invb - 1.0) * Dxx + invb * S[SXVX][idx]) / rho;", 1)],
    # stress_fluid: one memory component dropped from the pressure rate
    "fluid_p_memory_term": [("+ invb * xx * S[FXEX][idx] + invb * xz * S[FXEZ][idx]);",
                             "+ invb * xx * S[FXEX][idx]);", 1)],
    # stress_solid bottom strip: the wrong metric ratio in the zeta projection
    "solid_bottom_projection": [("const double Dvx = dvxz + (gxize / gzeze) * dvxx;\n"
                                 "                    const double Dvz = dvzz + (gxize / gzeze) * dvzx;\n"
                                 "                    const double be = o.be_zs[iz]",
                                 "const double Dvx = dvxz + (gxize / gxixi) * dvxx;\n"
                                 "                    const double Dvz = dvzz + (gxize / gzeze) * dvzx;\n"
                                 "                    const double be = o.be_zs[iz]", 1)],
    # velocity_solid bottom strip: xi metric in the zeta projection
    "solid_v_bottom_projection": [("Dzz = zx * (gxize / gzeze * e + f) + zz * (gxize / gzeze * c + d);",
                                   "Dzz = zx * (gxize / gzeze * e + f) + xz * (gxize / gzeze * c + d);", 1)],
    # velocity_fluid conservative rows: ze_x for xi_z in the xi projection
    "fluid_cons_projection": [("Dx = xx * (xx * sx + xz * sz) / (gxixi * J);",
                               "Dx = xx * (xx * sx + zx * sz) / (gxixi * J);", 1)],
    # solid zeta flux: the two metric factors exchanged
    "solid_zeta_flux": [("fx = o.J[g] * (o.zex[g] * S[SSX][i] + o.zez[g] * S[SSXZ][i]);",
                         "fx = o.J[g] * (o.zez[g] * S[SSX][i] + o.zex[g] * S[SSXZ][i]);", 1)],
    # fluid ghost flux: the odd extension about the interface replaced by an even one
    "fluid_ghost_even": [("o.Gfx[(n - 1) * nx + ix] = 2 * f0x - fx;", "o.Gfx[(n - 1) * nx + ix] = fx;", 1)],
    # interface condition: the tangential fluid velocity rebuilt with the wrong sign
    "iface_tangential_sign": [("Kd[FVX][f] = n1 * vnf + n2 * vtf;", "Kd[FVX][f] = n1 * vnf - n2 * vtf;", 1)],
    # interface condition: sigma_nt not cleared in the F direction
    "iface_snt_kept": [("if (dir == F) { vnf = vns; pf = -snn; snt = 0; }", "if (dir == F) { vnf = vns; pf = -snn; }", 1)],
    # filter: the pressure's sign in the averaged normal stress
    "filter_pressure_sign": [("sn = 0.5 * (snn - v[2]);", "sn = 0.5 * (snn + v[2]);", 1)],
    # filter: the solid takes the fluid's tangential velocity (no slip)
    "filter_no_slip": [("out[3 * nx + ix] = n1 * vn + n2 * vts;\n        out[4 * nx + ix] = n2 * vn - n1 * vts;",
                        "out[3 * nx + ix] = n1 * vn + n2 * vtf;\n        out[4 * nx + ix] = n2 * vn - n1 * vtf;", 1)],
}


def _build(name, src, outdir):
    text = src
    for old, new, count in MUTANTS[name]:
        assert text.count(old) >= count, (name, old)
        text = text.replace(old, new, count)
    cpp = os.path.join(outdir, f"{name}.cpp")
    so = os.path.join(outdir, f"{name}.so")
    with open(cpp, "w") as f:
        f.write(text)
    r = subprocess.run(["g++", "-O2", "-march=x86-64-v2", "-std=c++17", "-fPIC", "-fopenmp", "-shared", cpp, "-o", so],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return so


def _pins(so, threads):
    env = dict(os.environ, MSFDM_ORACLE_LIB=so, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *PINS],
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=900)
    failed = [ln.split()[1] for ln in r.stdout.splitlines() if ln.startswith("FAILED")]
    return r.returncode, failed, r.stdout[-3000:]


@pytest.fixture(scope="module")
def mutant_results(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    outdir = str(tmp_path_factory.mktemp("oracle_mutants"))
    with open(ORACLE) as f:
        src = f.read()
    workers = max(1, min(8, os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        libs = dict(zip(MUTANTS, ex.map(lambda n: _build(n, src, outdir), MUTANTS)))
    jobs = max(1, workers // 2)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        res = dict(zip(libs, ex.map(lambda n: _pins(libs[n], 2), libs)))
    return res


def test_unmutated_copy_passes_every_pin(mutant_results):
    rc, failed, tail = mutant_results["unmutated"]
    assert rc == 0 and not failed, tail


@pytest.mark.parametrize("name", sorted(n for n in MUTANTS if n != "unmutated"))
def test_every_planted_mistake_fails_a_pin(mutant_results, name):
    rc, failed, tail = mutant_results[name]
    assert rc == 1 and failed, f"mutant {name} passed every pin (rc {rc}):\n{tail}"
