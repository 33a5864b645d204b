"""The oracle pins catch planted mistakes.

Each case copies the oracle's C source, plants one plausible transcription
mistake (a flipped sign, a dropped term, swapped components or psi families,
a wrong exponent or index), builds the copy, and runs the pin suites against
it in a child process (SPARC_ORACLE_LIB points the oracle's loader at the
copy).  The pins must fail for every mutant.  CPU only: the mistakes are
planted in the oracle, never in GPU code.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PINS = ["tests/test_oracle_rhs_pins.py", "tests/test_oracle_pins.py"]

# (name, original text, planted text) in oracle/sparc_oracle.c
MUTANTS = [
    ("gravity sign", "(gpz[i] + rho[i] * c->g[k]) - vz[i] * sp;", "(gpz[i] - rho[i] * c->g[k]) - vz[i] * sp;"),
    ("rho0 in place of rho in the buoyancy", "(gpz[i] + rho[i] * c->g[k]) - vz[i] * sp;",
     "(gpz[i] + c->rho0[i] * c->g[k]) - vz[i] * sp;"),
    ("dropped vz d_z rho0", "s_rho[i] = -c->gradrho0_z[i] * vz[i] - c->rho0[i] * div",
     "s_rho[i] = - c->rho0[i] * div"),
    ("vz d_z p0 sign", "- vz[i] * c->gradp0_z[i] - sp * p[i];", "+ vz[i] * c->gradp0_z[i] - sp * p[i];"),
    ("induction sign", "double rbx = red * (dyf3[i] - dzf2[i]);", "double rbx = red * (dyf3[i] + dzf2[i]);"),
    ("curl b component sign", "double cby = -dxbz[i] + dzbx[i];", "double cby = dxbz[i] - dzbx[i];"),
    ("limiter dropped from Lorentz", "double rvy = -ri * (gpy[i] - (c2 * red)) - sp * vy[i];",
     "double rvy = -ri * (gpy[i] - (c2)) - sp * vy[i];"),
    ("background-current term sign", "rvx = rvx + ri * e1 * red;", "rvx = rvx - ri * e1 * red;"),
    ("grad rho0 components swapped", "-c->gradrho0_x[i] * vx[i] - c->gradrho0_y[i] * vy[i] -",
     "-c->gradrho0_x[i] * vy[i] - c->gradrho0_y[i] * vx[i] -"),
    ("v.grad p0 sign", "- vy[i] * c->gradp0_y[i] -", "+ vy[i] * c->gradp0_y[i] -"),
    ("CPML kappa multiplied", "d[gp] = d[gp] / c->pk[f][fp] + psi[fp];", "d[gp] = d[gp] * c->pk[f][fp] + psi[fp];"),
    ("CPML psi feedback sign", "d[gp] = d[gp] / c->pk[f][fp] + psi[fp];", "d[gp] = d[gp] / c->pk[f][fp] - psi[fp];"),
    ("CPML b psi sign", "rps[fp] = c->pa[f][fp] * d[gp] + c->pb[f][fp] * psi[fp];",
     "rps[fp] = c->pa[f][fp] * d[gp] - c->pb[f][fp] * psi[fp];"),
    ("z families 2/3 swapped", "    pml_apply(c, f, 2, dzbx);\n    pml_apply(c, f, 3, dzby);",
     "    pml_apply(c, f, 3, dzbx);\n    pml_apply(c, f, 2, dzby);"),
    ("x families 4/5 swapped", "    pml_apply(c, f, 4, dxf3);\n    pml_apply(c, f, 5, dxf2);",
     "    pml_apply(c, f, 5, dxf3);\n    pml_apply(c, f, 4, dxf2);"),
    ("curl_z B0 sign", "c->curlboz[p] = c->curlboz[p] - t[p];", "c->curlboz[p] = c->curlboz[p] + t[p];"),
    ("x/y kappa grading exponent", "double kap = 1.0 + (pc[2] - 1.0) * pow(s, pc[1]);\n              c->pk[f][fp] = kap;",
     "double kap = 1.0 + (pc[2] - 1.0) * pow(s, pc[1] + 1.0);\n              c->pk[f][fp] = kap;"),
    ("x alpha reversed", ": pml_f0 * (c->xaxis[ig] - c->xaxis[0]) / L;",
     ": pml_f0 * (c->xaxis[npl - 1] - c->xaxis[ig]) / L;"),
    ("y smoothing neighbour", "a[IX(c, i, j, k)] = (tmp[IX(c, i, j - 1, k)] + tmp[IX(c, i, j + 1, k)]) / 2.0;",
     "a[IX(c, i, j, k)] = (tmp[IX(c, i, j, k)] + tmp[IX(c, i, j + 1, k)]) / 2.0;"),
    ("z smoothing range", "      for (int k = 1; k <= gz - 1; ++k)", "      for (int k = 1; k <= gz - 2; ++k)"),
    ("psi stage weight", "c->ppsi[f][i] = c->pvars[f][i] + b * c->pscr[f][i];",
     "c->ppsi[f][i] = c->pvars[f][i] + 0.5 * b * c->pscr[f][i];"),
]


def _run_mutant(tmp, idx, name, old, new, src):
    d = os.path.join(tmp, f"m{idx}")
    os.makedirs(d)
    for f in ("sparc_oracle.h", "Makefile"):
        shutil.copy(os.path.join(ROOT, "oracle", f), d)
    assert src.count(old) >= 1, f"{name}: pattern not found in the oracle source"
    with open(os.path.join(d, "sparc_oracle.c"), "w") as fh:
        fh.write(src.replace(old, new))
    subprocess.run(["make", "-s", "-C", d, "liboracle.so"], check=True, capture_output=True)
    env = dict(os.environ, SPARC_ORACLE_LIB=os.path.join(d, "liboracle.so"), OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *PINS],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    return name, r.returncode, (r.stdout.strip().splitlines() or [""])[-1]


def test_every_planted_mistake_fails_a_pin(tmp_path):
    src = open(os.path.join(ROOT, "oracle", "sparc_oracle.c")).read()
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        res = list(ex.map(lambda a: _run_mutant(str(tmp_path), *a, src),
                          [(i, *m) for i, m in enumerate(MUTANTS)]))
    survived = [(n, tail) for n, rc, tail in res if rc == 0]
    for n, rc, tail in res:
        print(f"{n:40s} rc={rc} {tail}")
    assert not survived, f"mutants the pins did not catch: {survived}"
