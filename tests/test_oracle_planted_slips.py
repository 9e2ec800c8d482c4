"""The oracle pins catch plausible mistakes (CPU only).

Each case plants one slip in a copy of oracle/lbmpsm_oracle.c (a sign, a
coefficient, a factor, an operand), builds that copy into a temporary
directory, and runs the pins that should catch it in a child process with
the planted library loaded in place of the real one.  The case passes when
the pins fail.  The real oracle is never modified.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

# (id, function the slip is planted in, original text, planted text, pins)
SLIPS = [
    ("zh3-vel-zp-transverse-sign", "static void zh3_vel_zp",
     "f[15] = GF(16) + rho_tmp * oneOverSix * (-uz + ux) - Nxz;",
     "f[15] = GF(16) + rho_tmp * oneOverSix * (-uz + ux) + Nxz;",
     ["tests/test_oracle_zou_he.py"]),
    ("zh2-dens-xp-normal-coefficient", "static void zh2_dens_xp",
     "f[3] = GF(1) - twoOverThree * rho_bc * ux_tmp;",
     "f[3] = GF(1) - oneOverThree * rho_bc * ux_tmp;",
     ["tests/test_oracle_zou_he.py"]),
    ("zh3-vel-xp-density-sign", "static void zh3_vel_xp",
     "double rho_tmp = 1.0 / (1.0 + ux)", "double rho_tmp = 1.0 / (1.0 - ux)",
     ["tests/test_oracle_zou_he.py"]),
    ("zh3-vel-yn-index", "static void zh3_vel_yn",
     "f[11] = GF(12) + rho_tmp * oneOverSix * (uy + uz) - Nzy;",
     "f[11] = GF(18) + rho_tmp * oneOverSix * (uy + uz) - Nzy;",
     ["tests/test_oracle_zou_he.py"]),
    ("psm-omega-for-one-minus-omega", "static inline void collide_one",
     "solid_coll1 = f0_solid1 - fi + oneMinusOmega * (fi - fi_eq);",
     "solid_coll1 = f0_solid1 - fi + omega * (fi - fi_eq);",
     ["tests/test_oracle_pins.py", "-k", "darcy"]),
    ("psm-fluid-part-without-1-minus-B", "static inline void collide_one",
     "fi + (1.0 - Btot) * (fi_eq - fi) * omega", "fi + (1.0) * (fi_eq - fi) * omega",
     ["tests/test_oracle_pins.py", "-k", "darcy"]),
    ("psm-solid-equilibrium-at-fluid-velocity", "static inline void collide_one",
     "double f0_solid2 = feq_vel(L, iq, rho, L->us + n1 * 6 + 3);",
     "double f0_solid2 = fi_eq;",
     ["tests/test_oracle_pins.py", "-k", "darcy"]),
    ("psm-renormalisation-dropped", "static inline void collide_one",
     "    B1 = B1 / Btot;\n    B2 = B2 / Btot;\n", "",
     ["tests/test_oracle_pins.py", "-k", "darcy"]),
]


def _plant(src, anchor, old, new):
    start = src.index(anchor)
    end = src.find("\n}\n", start)
    body = src[start:end]
    assert old in body, f"{old!r} not in {anchor}"
    return src[:start] + body.replace(old, new, 1) + src[end:]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("slip", SLIPS, ids=[s[0] for s in SLIPS])
def test_planted_slip_fails_a_pin(tmp_path, slip):
    _, anchor, old, new, pins = slip
    src = open(os.path.join(ORACLE, "lbmpsm_oracle.c")).read()
    (tmp_path / "lbmpsm_oracle.c").write_text(_plant(src, anchor, old, new))
    for f in ("lbmpsm_oracle.h", "lbmpsm_oracle.py"):
        shutil.copy(os.path.join(ORACLE, f), tmp_path / f)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-fopenmp", "-o",
                    str(tmp_path / "liblbmpsm_oracle.so"), str(tmp_path / "lbmpsm_oracle.c"),
                    "-lm"], check=True)
    code = ("import sys; sys.path.insert(0, %r); import lbmpsm_oracle; "
            "assert lbmpsm_oracle.LIB_PATH.startswith(%r); import pytest; "
            "sys.exit(pytest.main(['-q', '-x', '-p', 'no:cacheprovider'] + %r))"
            % (str(tmp_path), str(tmp_path), pins))
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert "failed" in r.stdout
