"""Each pin of tests/oracle_pins.py fails on a copy of the oracle with a
plausible mistake planted in the function it pins (a swapped constant, a
reversed comparison, a wrong sign or neighbour), so the pins are known to
discriminate. The copies are built from oracle/oracle.cpp with one string
replaced, into a temporary directory; the oracle itself is never changed."""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import oracle_pins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "oracle.cpp")

# (pin, planted mistake, text in oracle.cpp, replacement)
MUTANTS = [
    ("dt", "convective radius without the 1.37 factor", "conv *= 1.37 * g;", "conv *= g;"),
    ("dt", "geometry factor n_f/3 instead of 3/n_f", "const double g = 3.0 / nf;", "const double g = nf / 3.0;"),
    ("dt", "boundary offset sqrt(V) instead of 2 sqrt(V)",
     "s[0] = 2.0 * std::pow(m->cell_volume[cl], 1.0 / ND);", "s[0] = 1.0 * std::pow(m->cell_volume[cl], 1.0 / ND);"),
    ("dt", "cell 0 as the neighbour across a boundary face", "                    cr = cl;\n",
     "                    cr = 0;\n"),
    ("dt", "centroid distance instead of its normal component", "const double dxn = std::fabs(dot2(s, nu));",
     "const double dxn = std::sqrt(dot2(s, s));"),
    ("dt", "1.37 inside the acoustic square root", "acou = 1.37 * std::sqrt(g * acou);",
     "acou = std::sqrt(1.37 * g * acou);"),
    ("eno", "central weight dropped in the reference ENO branch",
     "if (cfg.teno_variant != ORC_TENO_REFERENCE) sw[v] = 0.0;", "sw[v] = 0.0;"),
    ("eno", "failing directional weights kept in the corrected variant",
     "else if (cfg.teno_variant != ORC_TENO_REFERENCE) sw[sl * NC + v] = 0.0;",
     "else if (false) sw[sl * NC + v] = 0.0;"),
    ("eno", "passing directional weight 1 instead of 1/K", "sw[sl * NC + v] = (1.0 / K);", "sw[sl * NC + v] = 1.0;"),
    ("eno", "central weight in the directional normalisation",
     "                    sd = 0.0;\n                    for (int sl = 1; sl < gsz; ++sl) sd += sw[sl * NC + v];",
     "                    sd = 0.0;\n                    for (int sl = 0; sl < gsz; ++sl) sd += sw[sl * NC + v];"),
    ("eno", "face offset ipt/ta_0 dropped from the reference evaluation",
     "(b2(poly[2 * k], poly[2 * k + 1], xq[0], xq[1]) + (ipt[k] / ta[soff]))",
     "(b2(poly[2 * k], poly[2 * k + 1], xq[0], xq[1]))"),
    ("directional", "cone test with the wrong sign", "for (int i = 0; i < ND; i++) if (dt[i] < 0.0) in = false;",
     "for (int i = 0; i < ND; i++) if (dt[i] > 0.0) in = false;"),
    ("directional", "outward fill grown from the target cell alone", "rings.push_back(S[s]);",
     "rings.push_back(std::vector<uint32_t>{c});"),
    ("boundary", "wall pressure flux with the wrong sign", "Ft[1] = Wl[2] * nu[0]; Ft[2] = Wl[2] * nu[1];",
     "Ft[1] = -Wl[2] * nu[0]; Ft[2] = -Wl[2] * nu[1];"),
    ("boundary", "fixed-state density without the gas constant", "const double rho_bc = bc->p / (phys.R * bc->T);",
     "const double rho_bc = bc->p / bc->T;"),
    ("boundary", "fixed-state velocity components swapped", "Wr[0] = bc->u[0]; Wr[1] = bc->u[1];",
     "Wr[0] = bc->u[1]; Wr[1] = bc->u[0];"),
    ("boundary", "p_out subsonic test reversed", "const double pout = (um < a) ? bc->p : pl;",
     "const double pout = (um > a) ? bc->p : pl;"),
    ("boundary", "p_out temperature from the back pressure", "const double Tb = Tl;",
     "const double Tb = pout / (phys.R * rl);"),
    ("hll", "wave speeds swapped in the flux numerator",
     "F[i] = (Sr * s.Fl[i] - Sl * s.Fr[i] + Sl * Sr * (s.Ur[i] - s.Ul[i])) / (Sr - Sl);",
     "F[i] = (Sl * s.Fl[i] - Sr * s.Fr[i] + Sl * Sr * (s.Ur[i] - s.Ul[i])) / (Sr - Sl);"),
    ("hll", "jump term with the wrong sign",
     "F[i] = (Sr * s.Fl[i] - Sl * s.Fr[i] + Sl * Sr * (s.Ur[i] - s.Ul[i])) / (Sr - Sl);",
     "F[i] = (Sr * s.Fl[i] - Sl * s.Fr[i] + Sl * Sr * (s.Ul[i] - s.Ur[i])) / (Sr - Sl);"),
    ("hll", "two-shock estimate with the velocity jump reversed",
     "ps = (gl0 * pl + gr0 * pr - (ur - ul)) / (gl0 + gr0);", "ps = (gl0 * pl + gr0 * pr - (ul - ur)) / (gl0 + gr0);"),
    ("positivity", "density bound without the 1e-13 margin",
     "theta = std::fmin(theta, (B0[0] - er) / (B0[0] - B1[0]));", "theta = std::fmin(theta, B0[0] / (B0[0] - B1[0]));"),
    ("positivity", "pressure bisection returns its failing end", "            theta = lo;\n", "            theta = hi;\n"),
    ("positivity", "interior low-order flux from side 0 alone", "            phys.prims(Wr, U1);\n            rr = U1[0];",
     "            phys.prims(Wr, U0);\n            rr = U0[0];"),
    ("positivity", "low-order bracket not checked", "if (!(B0[0] > 0.0) || !(p0 > 0.0) || !std::isfinite",
     "if (false || !std::isfinite"),
    ("hll", "linearised estimate with the velocity jump reversed",
     "ps = 0.5 * (pl + pr) + 0.5 * (ul - ur) * ravg * aavg;", "ps = 0.5 * (pl + pr) + 0.5 * (ur - ul) * ravg * aavg;"),
    ("roe", "arithmetic mean instead of the Roe average of u",
     "const double u = (sl * ul[0] + sr * ur[0]) * isum;", "const double u = 0.5 * (ul[0] + ur[0]);"),
    ("roe", "arithmetic mean instead of the Roe average of H",
     "const double H = (sl * Hl + sr * Hr) * isum;", "const double H = 0.5 * (Hl + Hr);"),
    ("roe", "first wave strength with the wrong sign", "const double al1 = (dp - rho * a * dun) / (2.0 * a2);",
     "const double al1 = (dp + rho * a * dun) / (2.0 * a2);"),
    ("roe", "first eigenvector's energy with the wrong sign",
     "const double r1[4] = {1.0, u - a * n[0], v - a * n[1], H - a * un};",
     "const double r1[4] = {1.0, u - a * n[0], v - a * n[1], H + a * un};"),
]


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    src = open(SRC).read()
    out = tmp_path_factory.mktemp("oracle_mutants")

    def build(i):
        pin, what, old, new = MUTANTS[i]
        assert src.count(old) == 1, (what, src.count(old))
        cpp = out / ("m%02d.cpp" % i)
        so = out / ("m%02d.so" % i)
        cpp.write_text(src.replace(old, new))
        subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                        "-I", os.path.join(ROOT, "oracle"), "-shared", "-o", str(so), str(cpp)], check=True)
        return str(so)

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        return list(ex.map(build, range(len(MUTANTS))))


@pytest.mark.parametrize("i", range(len(MUTANTS)), ids=["%s: %s" % (m[0], m[1]) for m in MUTANTS])
def test_pin_fails_on_planted_mistake(oracle, mutant_libs, i):
    pin, what, _, _ = MUTANTS[i]
    with oracle.use_library(mutant_libs[i]):
        with pytest.raises((AssertionError, RuntimeError)):
            oracle_pins.CHECKS[pin](oracle)


def test_every_pin_has_mutants():
    assert {m[0] for m in MUTANTS} == set(oracle_pins.CHECKS)
