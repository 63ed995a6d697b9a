"""Mutation check of the oracle's pins — CPU only.

Each entry below plants one plausible transcription mistake in a copy of
oracle/vof_oracle.cpp (a dropped term, a wrong sign, index, upwind side or
axis, a swapped operand), builds that copy into a temporary library and runs
the oracle's pin suites (tests/test_oracle.py, tests/test_oracle_pins.py)
against it through VOF_ORACLE_LIB.  Every mutant must make the pins fail:
a mistake the pins let through would mean the function it sits in is not
pinned.  The golden-trajectory tests are deliberately not used here — they
were written by the oracle itself and would flag any change, pinned or not.

The mistakes are planted on the CPU only, in copies under a temporary
directory; the oracle source in the tree is never modified.
"""
import concurrent.futures as cf
import os
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "vof_oracle.cpp")
PIN_SUITES = [os.path.join(ROOT, "tests", "test_oracle.py"), os.path.join(ROOT, "tests", "test_oracle_pins.py")]

# (name, old text, new text): old must occur exactly once in the oracle source.
MUTANTS = [
    # helpers, IC, rho/nu (voftrial1.cu:52-104)
    ("clamp01 as a plain clamp", "return a + b + c - mx - mn;", "return std::fmin(std::fmax(c, a), b);"),
    ("IC predicate strict on x", "(x >= 0.0f && x <= fLx / 3.0f)", "(x >= 0.0f && x < fLx / 3.0f)"),
    ("nu phases swapped", "nu[t] = nu_l * f + nu_g * (R(1) - f);", "nu[t] = nu_g * f + nu_l * (R(1) - f);"),
    # predictor u face (:230-285)
    ("advect_u: y upwind side reversed",
     "const R ddy = (vh > R(0)) ? (uc - uym) * dyi : (uyp - uc) * dyi;",
     "const R ddy = (vh < R(0)) ? (uc - uym) * dyi : (uyp - uc) * dyi;"),
    # predictor v face (:287-342)
    ("advect_v: w average on the wrong row",
     "const R wh = R(0.25) * (w[id(i, j - 1, k + 1)] + w[id(i, j - 1, k)] +",
     "const R wh = R(0.25) * (w[id(i, j + 1, k + 1)] + w[id(i, j + 1, k)] +"),
    ("advect_v: u average on the wrong column",
     "const R uh = R(0.25) * (u[id(i, j - 1, k)] + u[id(i, j, k)] +\n"
     "                                            u[id(i + 1, j - 1, k)] + u[id(i + 1, j, k)]);",
     "const R uh = R(0.25) * (u[id(i, j - 1, k)] + u[id(i, j, k)] +\n"
     "                                            u[id(i - 1, j - 1, k)] + u[id(i - 1, j, k)]);"),
    ("advect_v: self-advection upwind side reversed",
     "const R ddy = (vc > R(0)) ? (vc - vym) * dyi : (vyp - vc) * dyi;",
     "const R ddy = (vc < R(0)) ? (vc - vym) * dyi : (vyp - vc) * dyi;"),
    ("advect_v: z viscous term dropped",
     "n0 * (vym - R(2) * vc + vyp) * dyi * dyi +\n                                                  n0 * (vzm - R(2) * vc + vzp) * dzi * dzi;",
     "n0 * (vym - R(2) * vc + vyp) * dyi * dyi;"),
    ("advect_v: gravity component", "vs[c] = vc + dt * (visc + adv + gy);", "vs[c] = vc + dt * (visc + adv + gx);"),
    ("advect_v: x advection by the wrong velocity",
     "const R adv = -uh * ddx - vc * ddy - wh * ddz;", "const R adv = -wh * ddx - vc * ddy - uh * ddz;"),
    # predictor w face (:344-399)
    ("advect_w: u average one plane off",
     "const R uh = R(0.25) * (u[id(i + 1, j, k - 1)] + u[id(i, j, k - 1)] +",
     "const R uh = R(0.25) * (u[id(i + 1, j, k + 1)] + u[id(i, j, k + 1)] +"),
    ("advect_w: v average on the wrong row",
     "const R vh = R(0.25) * (v[id(i, j + 1, k - 1)] + v[id(i, j, k - 1)] +\n"
     "                                            v[id(i, j, k)] + v[id(i, j + 1, k)]);",
     "const R vh = R(0.25) * (v[id(i, j - 1, k - 1)] + v[id(i, j, k - 1)] +\n"
     "                                            v[id(i, j, k)] + v[id(i, j - 1, k)]);"),
    ("advect_w: x upwind side reversed",
     "const R ddx = (uh > R(0)) ? (wc - wxm) * dxi : (wxp - wc) * dxi;",
     "const R ddx = (uh < R(0)) ? (wc - wxm) * dxi : (wxp - wc) * dxi;"),
    ("advect_w: y viscous term with dxi",
     "n0 * (wym - R(2) * wc + wyp) * dyi * dyi +", "n0 * (wym - R(2) * wc + wyp) * dxi * dxi +"),
    # Jacobi right-hand side and sweep (:401-439)
    ("rhs: v* difference on the wrong face",
     "(vs[id(i, j + 1, k)] - vs[c]) * dyi", "(vs[c] - vs[id(i, j - 1, k)]) * dyi"),
    ("rhs: w* difference reversed", "(ws[id(i, j, k + 1)] - ws[c]) * dzi", "(ws[c] - ws[id(i, j, k + 1)]) * dzi"),
    ("rhs: w* term with dyi", "(ws[id(i, j, k + 1)] - ws[c]) * dzi", "(ws[id(i, j, k + 1)] - ws[c]) * dyi"),
    ("jacobi: north coefficient tested on the wrong wall",
     "const R an = (j != jmax) ? (dyi * dyi) : R(0);", "const R an = (j != jmin) ? (dyi * dyi) : R(0);"),
    ("jacobi: back coefficient with dyi",
     "const R ab = (k != kmin) ? (dzi * dzi) : R(0);", "const R ab = (k != kmin) ? (dyi * dyi) : R(0);"),
    ("jacobi: front neighbour on the wrong side", "af * p[id(i, j, k + 1)]", "af * p[id(i, j, k - 1)]"),
    # correction (:441-493)
    ("update_v: pressure gradient with dxi",
     "v[c] = vs[c] - dt / r * (p[c] - p[m]) * dyi;", "v[c] = vs[c] - dt / r * (p[c] - p[m]) * dxi;"),
    ("update_w: density average on the wrong neighbour",
     "const size_t c = id(i, j, k), m = id(i, j, k - 1);\n                    const R r = R(0.5) * (rho[c] + rho[m]);\n"
     "                    w[c] = ws[c] - dt / r * (p[c] - p[m]) * dzi;",
     "const size_t c = id(i, j, k), m = id(i, j, k - 1);\n                    const R r = R(0.5) * (rho[c] + rho[id(i, j - 1, k)]);\n"
     "                    w[c] = ws[c] - dt / r * (p[c] - p[m]) * dzi;"),
    ("update_w: gradient sign", "w[c] = ws[c] - dt / r * (p[c] - p[m]) * dzi;", "w[c] = ws[c] + dt / r * (p[c] - p[m]) * dzi;"),
    # velocity BC (:959-1051, :1107-1120)
    ("bc_vel: v wall not zeroed at the top", "u[gt] = t * u[it]; v[gt] = R(0); w[gt] = t * w[it];",
     "u[gt] = t * u[it]; v[gt] = t * v[it]; w[gt] = t * w[it];"),
    ("bc_vel: u ghost negated on the k pass in free slip",
     "u[gb] = t * u[ib]; v[gb] = t * v[ib]; w[ib] = R(0);", "u[gb] = -t * u[ib]; v[gb] = t * v[ib]; w[ib] = R(0);"),
    ("bc_vel: no-slip sign dropped on the j pass",
     "u[gt] = t * u[it]; v[gt] = R(0); w[gt] = t * w[it];", "u[gt] = u[it]; v[gt] = R(0); w[gt] = t * w[it];"),
    ("bc_vel: no-slip ghosts zeroed instead of negated",
     "const R t = (flags & VOF_FLAG_NO_SLIP) ? R(-1) : R(1);", "const R t = (flags & VOF_FLAG_NO_SLIP) ? R(0) : R(1);"),
    # density-weighted viscous term (extension)
    ("rho visc: face density of one cell", "const R rf = R(0.5) * (a[O_RHO][ca] + a[O_RHO][cb]);",
     "const R rf = a[O_RHO][cb];"),
    ("rho visc: upper edge averaged on the lower cells",
     "mp = R(0.25) * ((ma + mb) + (mu_at(ca + o) + mu_at(cb + o)));",
     "mp = R(0.25) * ((ma + mb) + (mu_at(ca - o) + mu_at(cb - o)));"),
    ("rho visc: lower edge viscosity of the face's cells only",
     "mm = R(0.25) * ((mu_at(ca - o) + mu_at(cb - o)) + (ma + mb));", "mm = R(0.5) * (ma + mb);"),
    ("rho visc: mu is nu (density dropped)", "inline R mu_at(size_t q) const { return a[O_RHO][q] * a[O_NU][q]; }",
     "inline R mu_at(size_t q) const { return a[O_NU][q]; }"),
    ("rho visc: own-axis viscosities swapped",
     "                mp = mb;\n                mm = ma;", "                mp = ma;\n                mm = mb;"),
    # FCT (:509-905)
    ("fct: high-order flux central instead of downwind",
     "static inline R flux_high(R vel, R dt_, R fm, R fp) { return (vel <= R(0)) ? vel * dt_ * fm : vel * dt_ * fp; }",
     "static inline R flux_high(R vel, R dt_, R fm, R fp) { return R(0.5) * vel * dt_ * (fm + fp); }"),
    ("fct: high-order flux upwind (no anti-diffusion)",
     "static inline R flux_high(R vel, R dt_, R fm, R fp) { return (vel <= R(0)) ? vel * dt_ * fm : vel * dt_ * fp; }",
     "static inline R flux_high(R vel, R dt_, R fm, R fp) { return (vel >= R(0)) ? vel * dt_ * fm : vel * dt_ * fp; }"),
    ("fct: donor cell downwind",
     "static inline R flux_low(R vel, R dt_, R fm, R fp) { return (vel >= R(0)) ? vel * dt_ * fm : vel * dt_ * fp; }",
     "static inline R flux_low(R vel, R dt_, R fm, R fp) { return (vel <= R(0)) ? vel * dt_ * fm : vel * dt_ * fp; }"),
    ("fct_x: limiter side swapped",
     "cx[cxp] = (ax[cxp] >= R(0)) ? std::fmin(rp[cxp], rm[c]) : std::fmin(rp[c], rm[cxp]);\n"
     "                    cy[cyp] = (ay[cyp] >= R(0)) ? std::fmin(rp[cyp], rm[c]) : std::fmin(rp[c], rm[cyp]);\n"
     "                }\n        // stage 4 (:619-646)",
     "cx[cxp] = (ax[cxp] >= R(0)) ? std::fmin(rp[c], rm[cxp]) : std::fmin(rp[cxp], rm[c]);\n"
     "                    cy[cyp] = (ay[cyp] >= R(0)) ? std::fmin(rp[cyp], rm[c]) : std::fmin(rp[c], rm[cyp]);\n"
     "                }\n        // stage 4 (:619-646)"),
    ("fct_x: stage 4 without the dilation factor",
     "ay[c] * cy[c]) / dy) * vol / dv;\n                    if (stats_on) a[O_S4][c] = out;\n                    F[c] = clamp01(out);\n                }\n    }\n\n    // ================= FCT y",
     "ay[c] * cy[c]) / dy);\n                    if (stats_on) a[O_S4][c] = out;\n                    F[c] = clamp01(out);\n                }\n    }\n\n    // ================= FCT y"),
    ("fct_x: stage 1 dilation sign",
     "const R dv = vol - dt * dy * dz * (u[cp] - u[c]);", "const R dv = vol + dt * dy * dz * (u[cp] - u[c]);"),
    ("fct_x: Zalesak P+ from the wrong faces",
     "const R pp = std::fmax(R(0), ax[c]) - std::fmin(R(0), ax[cp]) +\n                                 std::fmax(R(0), ay[c]) - std::fmin(R(0), ay[cyp]);",
     "const R pp = std::fmax(R(0), ax[cp]) - std::fmin(R(0), ax[c]) +\n                                 std::fmax(R(0), ay[c]) - std::fmin(R(0), ay[cyp]);"),
    ("fct_y: low flux of the top face on the wrong cells",
     "const R ft = flux_low(v[cp], dt, F[c], F[cp]);", "const R ft = flux_low(v[cp], dt, F[cm], F[c]);"),
    ("fct_y: Q- without the spacing", "const R qm = (Ftd[c] - fmn) * dx;\n                    rm[c] = (pm > R(0)) ? std::fmin(R(1), qm / pm) : R(0);\n                }\n        // stage 3 (:727-756)",
     "const R qm = (Ftd[c] - fmn);\n                    rm[c] = (pm > R(0)) ? std::fmin(R(1), qm / pm) : R(0);\n                }\n        // stage 3 (:727-756)"),
    ("fct_z: anti-diffusive flux sign", "az[cp] = ffH - ffL;\n                    az[c] = fbH - fbL;",
     "az[cp] = ffL - ffH;\n                    az[c] = fbL - fbH;"),
    ("fct_z: limiter of the back face", "cz[czp] = (az[czp] >= R(0)) ? std::fmin(rp[czp], rm[c]) : std::fmin(rp[c], rm[czp]);",
     "cz[czp] = (az[czp] >= R(0)) ? std::fmin(rp[czp], rm[czp]) : std::fmin(rp[c], rm[c]);"),
    ("fct_z: stage 1 clip dropped", "if (out > R(1) || out < R(0)) out = clamp01(out);\n                    Ftd[c] = out;\n                }\n        // stage 2 (:816-859)",
     "Ftd[c] = out;\n                }\n        // stage 2 (:816-859)"),
    # F BC, clamp, diagnostics
    ("bc_F: top ghost copies the wrong row", "F[id(i, jmax + 1, k)] = F[id(i, jmax, k)];", "F[id(i, jmax + 1, k)] = F[id(i, jmax - 1, k)];"),
    ("diag: volume over a missing plane", "for (int k = kmin; k <= kmax; ++k)\n            for (int j = jmin; j <= jmax; ++j)\n                for (int i = imin; i <= imax; ++i) {\n                    const size_t c = id(i, j, k);\n                    vs += double(F[c]);",
     "for (int k = kmin; k < kmax; ++k)\n            for (int j = jmin; j <= jmax; ++j)\n                for (int i = imin; i <= imax; ++i) {\n                    const size_t c = id(i, j, k);\n                    vs += double(F[c]);"),
]


def _run_mutant(idx, tmp):
    name, old, new = MUTANTS[idx]
    src = open(SRC).read()
    n = src.count(old)
    if n != 1:
        return name, f"pattern occurs {n} times"
    d = os.path.join(tmp, f"m{idx}")
    os.makedirs(d)
    path = os.path.join(d, "vof_oracle.cpp")
    with open(path, "w") as f:
        f.write(src.replace(old, new))
    lib = os.path.join(d, "liboracle.so")
    cc = subprocess.run(["g++", "-O1", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
                         path, "-o", lib], capture_output=True, text=True)
    if cc.returncode != 0:
        return name, "does not compile: " + cc.stderr[-400:]
    env = dict(os.environ, VOF_ORACLE_LIB=lib, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", *PIN_SUITES, "-x", "-q", "-p", "no:cacheprovider"],
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=600)
    return name, ("caught" if r.returncode == 1 else f"pins pass (rc {r.returncode})")


def test_every_planted_mistake_fails_a_pin():
    with tempfile.TemporaryDirectory() as tmp:
        # the source includes "../include/vof_capi.h": give each copy that path
        os.makedirs(os.path.join(tmp, "include"))
        with open(os.path.join(ROOT, "include", "vof_capi.h")) as f, \
                open(os.path.join(tmp, "include", "vof_capi.h"), "w") as g:
            g.write(f.read())
        with cf.ThreadPoolExecutor(max_workers=6) as ex:
            results = list(ex.map(lambda i: _run_mutant(i, tmp), range(len(MUTANTS))))
    missed = [(n, why) for n, why in results if why != "caught"]
    assert not missed, missed


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_pattern_is_unique(name, old, new):
    assert open(SRC).read().count(old) == 1, name
