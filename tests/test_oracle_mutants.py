"""Planted mistakes in the oracle must fail its pins (CPU only).

Each entry below copies oracle/mercury_oracle.c, replaces one exact fragment by
a plausible mistake (a dropped or halved term, a flipped sign, a transposed
component, a wrong coefficient or stencil), builds the copy into a temporary
library and runs the oracle's pin tests against it through MERCURY_ORACLE_LIB.
The pins are adequate only if every such build makes at least one of them
fail.  The fragment must occur exactly once in the source, so a mutant always
changes the line it names.
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "mercury_oracle.c")
PINS = ["tests/test_oracle_pins.py", "tests/test_oracle_term_pins.py", "tests/test_emf_pins.py", "tests/test_dipole_oracle.py",
        "tests/test_dec_oracle.py", "tests/test_species_oracle.py", "tests/test_singular_oracle.py"]

# (what the mistake is, exact fragment, replacement)
MUTANTS = [
    # singular edges (SingularEdgeRegistry.cpp:78-330, MagExplicit_Ideal.cpp:57-130,
    # MagExplicit_Hall.cpp:32-56, Current.cpp:59-138)
    ("singular: incidence sign flipped", "if (fa == 1 && f[0] == e[0] && f[1] == e[1]) return f[2] == e[2] - 1 ? 1 : (f[2] == e[2] ? -1 : 0);",
     "if (fa == 1 && f[0] == e[0] && f[1] == e[1]) return f[2] == e[2] - 1 ? -1 : (f[2] == e[2] ? 1 : 0);"),
    ("singular: dual area from volume / length", "        if (nc >= 3) {\n            double ctr[3]", "        if (nc >= 30) {\n            double ctr[3]"),
    ("singular: cell weights not normalised", "S->cw[n] = vsum > 0.0 ? cw[u] / vsum : 0.0", "S->cw[n] = cw[u]"),
    ("singular: face coefficient without inverse Hodge", "S->fcoef[n] = E.ih * fcoef_sign[w];", "S->fcoef[n] = fcoef_sign[w];"),
    ("singular: edge vector without orientation", "dr[q] += orient * F(bk, S_DR + 3 * ax + q, e[0], e[1], e[2]);",
     "dr[q] += F(bk, S_DR + 3 * ax + q, e[0], e[1], e[2]);"),
    ("singular: Hall taper dropped", "emf = taper * (ex * cv[0] + ey * cv[1] + ez * cv[2]);", "emf = (ex * cv[0] + ey * cv[1] + ez * cv[2]);"),
    ("singular: ideal EMF sign", "const double ex = -(u1 * b2 - u2 * b1), ey = -(u2 * b0 - u0 * b2), ez = -(u0 * b1 - u1 * b0);",
     "const double ex = (u1 * b2 - u2 * b1), ey = (u2 * b0 - u0 * b2), ez = (u0 * b1 - u1 * b0);"),
    ("singular: pressure difference halved twice", "const double dp = 0.5 * ((b->f[S_PVH + 3][q + d]", "const double dp = 0.25 * ((b->f[S_PVH + 3][q + d]"),
    ("singular: local resistivity decay", "emf += P->local_eta_max * exp(-rr / P->local_r_decay) * jedge;",
     "emf += P->local_eta_max * exp(-rr * P->local_r_decay) * jedge;"),
    # A7 Scheme_U_ Rusanov flux (MercurySolver_FluidRHS.cpp:32-102)
    ("flux: dissipation dropped", "- 0.5 * r * (Uc[1][m] - Uc[0][m])", "- 0.0 * r * (Uc[1][m] - Uc[0][m])"),
    ("flux: dissipation sign", "- 0.5 * r * (Uc[1][m] - Uc[0][m])", "- 0.5 * r * (Uc[0][m] - Uc[1][m])"),
    ("flux: dissipation halved", "- 0.5 * r * (Uc[1][m] - Uc[0][m])", "- 0.25 * r * (Uc[1][m] - Uc[0][m])"),
    ("flux: min of the two radii", "const double r = dmax(rad[0], rad[1]);", "const double r = dmin(rad[0], rad[1]);"),
    ("flux: left radius only", "const double r = dmax(rad[0], rad[1]);", "const double r = rad[0];"),
    ("flux: sound speed without |k|", "rad[side] = fabs(uvw) + c * sqrt(kk2);", "rad[side] = fabs(uvw) + c;"),
    ("flux: transposed area component", "Fl[side][2] = rho * uvw * v + k2 * p;", "Fl[side][2] = rho * uvw * v + k3 * p;"),
    ("flux: kinetic energy factor", "Uc[side][4] = p / (gm - 1.0) + 0.5 * rho * (u * u + v * v + w * w);",
     "Uc[side][4] = p / (gm - 1.0) + rho * (u * u + v * v + w * w);"),
    ("flux: divergence sign", "r -= (F(b, S_F + 5 + m, i, j + 1, k) - F(b, S_F + 5 + m, i, j, k)) * invJ;",
     "r += (F(b, S_F + 5 + m, i, j + 1, k) - F(b, S_F + 5 + m, i, j, k)) * invJ;"),
    # A9 AssembleOneDirectionEMF_ (MercurySolver_MagExplicit_Ideal.cpp:433-594)
    ("face EMF: dissipation on the total field", "return 0.5 * (GL + GR) - 0.5 * radius * (biR - biL);",
     "return 0.5 * (GL + GR) - 0.5 * radius * (btR - btL);"),
    ("face EMF: dissipation dropped", "return 0.5 * (GL + GR) - 0.5 * radius * (biR - biL);", "return 0.5 * (GL + GR);"),
    ("face EMF: radius from the component axis", "const double radius = dmax(fabs(ufL), fabs(ufR));",
     "const double radius = dmax(fabs(ubL), fabs(ubR));"),
    ("face EMF: background field dropped", "const double GL = ufL * btL - ubL * bff, GR = ufR * btR - ubR * bff;",
     "const double GL = ufL * biL - ubL * bff, GR = ufR * biR - ubR * bff;"),
    ("face EMF: contravariant velocity without 1/Jac", "const double inv_jac = 1.0 / (fabs(F(b, S_JAC, i, j, k)) + 1.0e-30);",
     "const double inv_jac = 1.0;"),
    # A10 AssembleEdgeEMF_FromFaceE_Ideal_ (MercurySolver_MagExplicit_Ideal.cpp:138-372)
    ("UCT: C_uct = 0.25", "const double Cuct = 0.5;", "const double Cuct = 0.25;"),
    ("UCT: C_uct = 1.0", "const double Cuct = 0.5;", "const double Cuct = 1.0;"),
    ("UCT: wrong speed axis", "const double ae = MAXU4(1, i, j, k, i, j - 1, k, i, j, k - 1, i, j - 1, k - 1);",
     "const double ae = MAXU4(2, i, j, k, i, j - 1, k, i, j, k - 1, i, j - 1, k - 1);"),
    ("UCT: jump sign", "const double dBz = F(b, S_B + 2, i, j, k) - F(b, S_B + 2, i, j - 1, k);",
     "const double dBz = F(b, S_B + 2, i, j - 1, k) - F(b, S_B + 2, i, j, k);"),
    ("UCT: face average over 2", "const double Ec = 0.25 * (F(b, S_EF + 2, i, j, k) + F(b, S_EF + 2, i, j - 1, k)",
     "const double Ec = 0.5 * (F(b, S_EF + 2, i, j, k) + F(b, S_EF + 2, i, j - 1, k)"),
    # A8 AddSourceToRHS_Fluid (MercurySolver_FluidRHS.cpp:206-527)
    ("sources: U+ dropped from the induction term", "const double subx = (v - upy) * Bz - (w - upz) * By;",
     "const double subx = v * Bz - w * By;"),
    ("sources: induction work dropped", "const double subu = subx * u + suby * v + subz * w;", "const double subu = 0.0;"),
    ("sources: lit/shadow frequencies swapped", "const double vst = in_shadow ? P->nu_shadow : P->nu_lit;",
     "const double vst = in_shadow ? P->nu_lit : P->nu_shadow;"),
    ("sources: shadow on the day side", "const int in_shadow = (x < 0.0 && rho_cyl < 1.0);",
     "const int in_shadow = (x > 0.0 && rho_cyl < 1.0);"),
    ("sources: partner mass", "const double Mo = P->M_Na;", "const double Mo = Ms;"),
    ("sources: thermal mass ratio", "const double b1 = vst * (P->Tn0 - Ts0) * (Ms / (Ms + Mo));",
     "const double b1 = vst * (P->Tn0 - Ts0) * (Mo / (Ms + Mo));"),
    ("sources: photo energy without Tn0", "F(b, sr + 4, i, j, k) += P->a6 * sss * P->Tn0;",
     "F(b, sr + 4, i, j, k) += P->a6 * sss;"),
    ("sources: photo mass coefficient", "F(b, sr + 0, i, j, k) += P->a1_Na * sss;", "F(b, sr + 0, i, j, k) += P->a6 * sss;"),
    ("sources: Hall weight n_s/n_e", "const double chi = num.w[s];", "const double chi = num.chi[s];"),
    ("sources: drag on the energy with u", "F(b, sr + 4, i, j, k) -= P->a4 * rho * u2 * vst;",
     "F(b, sr + 4, i, j, k) -= P->a4 * rho * u * vst;"),
    # A1 / A2 calc_PV, calc_Uplus, Hall_Num_Limiter (MercurySolver_Tools.cpp:111-301)
    ("U+: species weight", "double up = n.chi[0] * us[0][c] + n.chi[1] * us[1][c];",
     "double up = n.chi[0] * us[0][c] + n.chi[0] * us[1][c];"),
    ("limiter: floor dropped", "n.ne_eff = sqrt(n.ne * n.ne + P->ne_floor * P->ne_floor);", "n.ne_eff = n.ne;"),
    # A6 Compute_Timestep (MercurySolver_Timestep.cpp:8-141)
    ("time step: slow magnetosonic speed", "double cf = sqrt(0.5 * (t + sqrt(disc)));", "double cf = sqrt(0.5 * (t - sqrt(disc)));"),
    # A20 SetupCellReconstructionWeights_ (MercurySolver_ReconstructionSetup.cpp:29-191), Metric.cpp:505-553
    ("W_b: unweighted LSQ", "mu[n] = 1.0 / dmax(sqrt(dx * dx + dy * dy + dz * dz), 1.0e-14);", "mu[n] = 1.0;"),
    ("W_b: face order", "s[1][a] = F(b, S_JD + 0 + a, i + 1, j, k);", "s[1][a] = F(b, S_JD + 0 + a, i, j, k);"),
    ("eigen: one Jacobi sweep", "for (int it = 0; it < 18; ++it) {", "for (int it = 0; it < 1; ++it) {"),
    ("W_j: edge order", "{0, 0, 0, 1, 0, 0}, {0, 1, 0, 1, 1, 0}, {0, 0, 1, 1, 0, 1}",
     "{0, 0, 0, 1, 0, 0}, {0, 0, 1, 1, 0, 1}, {0, 1, 0, 1, 1, 0}"),
    ("alpha cap: ratio 1e-2", "const double RATIO = 1e-3, EABS = 1e-12;", "const double RATIO = 1e-2, EABS = 1e-12;"),
    ("alpha cap: ratio 1e-4", "const double RATIO = 1e-3, EABS = 1e-12;", "const double RATIO = 1e-4, EABS = 1e-12;"),
    # A11-A13 Hall, artificial and ambipolar edge EMFs
    ("Hall: C_wh = 0.2", "const double eps = 1e-14, Cwh = 0.1;", "const double eps = 1e-14, Cwh = 0.2;"),
    ("Hall: diffusion sign", "val = Ecen + 0.5 * mue * d1 - 0.5 * muz * d2;", "val = Ecen - 0.5 * mue * d1 - 0.5 * muz * d2;"),
    ("Hall: taper width", "tp = 0.5 * (1.0 + tanh(2.0 * (r - mid) / width));", "tp = 0.5 * (1.0 + tanh((r - mid) / width));"),
    ("ambipolar: density sum instead of mean", "for (int s = 0; s < ns; ++s) rs[s] = 0.25 * rs[s];",
     "for (int s = 0; s < ns; ++s) rs[s] = 1.0 * rs[s];"),
    ("ambipolar: node mean over 4", "pe[e] = 0.125 * sum;", "pe[e] = 0.25 * sum;"),
    ("artificial: linear ramp", "eta = P->arti_eta_max * (q * q * (3.0 - 2.0 * q));", "eta = P->arti_eta_max * q;"),
    ("local eta: squared distance", "eta = P->local_eta_max * exp(-rr / P->local_r_decay);",
     "eta = P->local_eta_max * exp(-rr * rr / P->local_r_decay);"),
    # A4 CurlAdjFaceToEdge, A16 CurlEdgeToFace
    ("adjoint curl: beta dropped", "#define BB(a, i, j, k) (F(b, S_BETA + (a), i, j, k) * F(b, sf + (a), i, j, k))",
     "#define BB(a, i, j, k) (F(b, sf + (a), i, j, k))"),
    # A20 DipoleField::Build_Badd_FaceFlux (Z4_Mercury/include/operators/Dipole.h:25-324)
    ("dipole: softening dropped", "double f = (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) + D[6] * D[6];",
     "double f = (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);"),
    ("dipole: one cross-product component flipped", "A[1] += (D[5] * r[0] - D[3] * r[2]) / f;",
     "A[1] += (D[3] * r[2] - D[5] * r[0]) / f;"),
    ("dipole: potential at the edge start", "xm[c] = (x1[c] + x2[c]) * 0.5;", "xm[c] = x1[c];"),
    ("dipole: power 1/2", "f = sqrt(f) * f;", "f = sqrt(f);"),
    ("dipole: curl of the eta face with the wrong sign",
     "else if (a == 1) c = (AE(0, i, j, k + 1) - AE(0, i, j, k)) + (AE(2, i, j, k) - AE(2, i + 1, j, k));",
     "else if (a == 1) c = (AE(0, i, j, k) - AE(0, i, j, k + 1)) + (AE(2, i, j, k) - AE(2, i + 1, j, k));"),
    # the third ion species, He++ (no reference counterpart; pinned by reductions to two species)
    ("He++: charge dropped from its charge density", "n.q[2] = P->Z_He * n.n[2];", "n.q[2] = n.n[2];"),
    ("He++: left out of n_e", "n.ne = n.ne + n.q[2];", "n.ne = n.ne;"),
    ("He++: pressure left out of p_e", "if (n_species(P) > 2) pe = pe + F(b, S_PVHE + 3, i, j, k);", "if (n_species(P) > 2) pe = pe;"),
    ("He++: dropped from U+", "if (ns > 2) up = up + n.chi[2] * us[2][c];", "if (ns > 2) up = up;"),
    ("induction force with the number density", "const double nsp = num.n[s], qs = num.q[s];",
     "const double nsp = num.n[s], qs = num.n[s];"),
    # src/8_dec DecExteriorDerivative.cpp / DecDiagnostics.cpp
    ("DEC d0: reversed difference", "for (int k = 0; k <= b.m[2]; ++k) ey[DQ(b, i, j, k)] = n[DQ(b, i, j + 1, k)] - n[DQ(b, i, j, k)];",
     "for (int k = 0; k <= b.m[2]; ++k) ey[DQ(b, i, j, k)] = n[DQ(b, i, j, k)] - n[DQ(b, i, j + 1, k)];"),
    ("DEC d1: eta face sign", "return (ex[DQ(b, i, j, k + 1)] - ex[DQ(b, i, j, k)]) + (ez[DQ(b, i, j, k)] - ez[DQ(b, i + 1, j, k)]);",
     "return (ex[DQ(b, i, j, k + 1)] - ex[DQ(b, i, j, k)]) + (ez[DQ(b, i + 1, j, k)] - ez[DQ(b, i, j, k)]);"),
    ("DEC d1: owned range of xi faces", "    for (int i = 0; i <= b.m[0]; ++i)\n        for (int j = 0; j < b.m[1]; ++j)\n            for (int k = 0; k < b.m[2]; ++k) {\n                const double v = dec_d1_xi",
     "    for (int i = 0; i < b.m[0]; ++i)\n        for (int j = 0; j < b.m[1]; ++j)\n            for (int k = 0; k < b.m[2]; ++k) {\n                const double v = dec_d1_xi"),
    ("DEC faraday: plus sign", "if (update) fz[DQ(b, i, j, k)] -= dt * v;", "if (update) fz[DQ(b, i, j, k)] += dt * v;"),
    ("DEC d2: zeta term dropped", "(fz[DQ(b, i, j, k + 1)] - fz[DQ(b, i, j, k)]);\n}", "0.0;\n}"),
]


def _build(tag, frag, repl, work):
    src = open(SRC).read()
    assert src.count(frag) == 1, f"{tag}: fragment occurs {src.count(frag)} times"
    d = os.path.join(work, f"m{abs(hash(tag)) % 10**8}")
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, "mutant.c")
    with open(path, "w") as f:
        f.write(src.replace(frag, repl))
    lib = os.path.join(d, "liboracle.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared", "-o", lib, path,
                    "-lm"], check=True, capture_output=True)
    return lib


def _run(lib):
    env = dict(os.environ, MERCURY_ORACLE_LIB=lib, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *PINS], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout[-400:]


def test_every_planted_mistake_fails_a_pin():
    with tempfile.TemporaryDirectory() as work:
        def one(m):
            tag, frag, repl = m
            rc, out = _run(_build(tag, frag, repl, work))
            return tag, rc, out

        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            results = list(ex.map(one, MUTANTS))
    survivors = [(t, out) for t, rc, out in results if rc == 0]
    broken = [(t, rc, out) for t, rc, out in results if rc not in (0, 1)]
    assert not broken, f"mutant runs that did not complete: {broken}"
    assert not survivors, "planted mistakes no pin catches: " + "; ".join(t for t, _ in survivors)


@pytest.mark.parametrize("tag,frag,repl", MUTANTS[:1])
def test_unmutated_build_passes(tag, frag, repl):
    """The harness itself: the source built the same way passes every pin."""
    with tempfile.TemporaryDirectory() as work:
        lib = _build("control", frag, frag, work)
        rc, out = _run(lib)
    assert rc == 0, out
