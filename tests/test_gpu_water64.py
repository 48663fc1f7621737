"""The large single system resident in HBM: (H2O)64 (seed 1, random
orientations, 1 g/cm^3, O-O >= 2.7 A) in cc-pVDZ / cc-pVDZ-JKFIT, nbf 1536,
naux 7424, 320 occupied orbitals.

The FP64 RI-RHF from the core guess converges to 1e-8 Eh; the mixed RI-RHF
runs integer builds until the DIIS error is below 1e-5 and then hands over to
an FP64 context for the end-game (src/SCF.cpp:78-223 as iterative
refinement), and lands within 1e-8 Eh of the FP64 run. On the converged
density the integer build stays within 1 uEh of the FP64 build. Each context
keeps one form of B (FP64 rows or integer forms), as a large system must;
the integer-only context fits in 70 GB."""
import sys
import time

import numpy as np
import pytest

from nopt_amd import geometry as G
from nopt_amd.jk import JKContext
from nopt_amd.molecule import Molecule
from nopt_amd.scf import rhf
from nopt_amd.workloads import two_electron_energy

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_water64_mixed_rhf_with_fp64_endgame():
    t0 = time.perf_counter()
    m = Molecule.from_angstrom(*G.water_cluster(64, 1), basis="cc-pvdz", aux="cc-pvdz-jkfit")
    assert (m.nbf, m.naux, m.nocc) == (1536, 7424, 320)
    S, T, V = m.one_electron()
    H = T + V
    metric = m.metric()
    eri3 = m.eri3_packed()
    say("(H2O)64 integrals", round(time.perf_counter() - t0, 1), "s")

    def context(keep):
        c = JKContext(m.nbf, m.naux, keep=keep)
        c.set_B(eri3, metric)
        say(keep, "context", round(c.device_bytes() / 1e9, 1), "GB on the device,", round(time.perf_counter() - t0, 1), "s")
        return c

    ctx = context("fp64")
    ref = rhf(m, ctx=ctx, H=H, S=S, prec="fp64", e_conv=1e-8, d_conv=1e-6)
    assert ref.converged
    Co = np.ascontiguousarray(ref.C[:, :m.nocc].T)
    D = Co.T @ Co
    J64, K64 = ctx.build(D, Co, m.nocc, prec="fp64")
    ctx.close()
    say("FP64 RI-RHF", ref.E, ref.builds, round(time.perf_counter() - t0, 1), "s")

    ctx = context("int8")
    t1 = time.perf_counter()
    J8, K8 = ctx.build(D, Co, m.nocc, prec="int8")
    t2 = time.perf_counter()
    J8, K8 = ctx.build(D, Co, m.nocc, prec="int8")
    dE = two_electron_energy(D, J8, K8) - two_electron_energy(D, J64, K64)
    dj, dk = float(np.abs(J8 - J64).max()), float(np.abs(K8 - K64).max())
    say("int8 vs fp64 at the converged density: dE", dE, "max|dJ|", dj, "max|dK|", dk, "; builds",
        round(t2 - t1, 3), round(time.perf_counter() - t2, 3), "s (host buffers);",
        round(ctx.device_bytes() / 1e9, 1), "GB on the device")
    assert abs(dE) < 1e-6 and dj < 1e-8 and dk < 5e-7
    # the integer-only context keeps the packed triangle alone, and the
    # exchange operand in digit planes per column group: no FP64 X
    assert ctx.device_bytes() < 70e9

    def to_fp64(c):
        c.close()
        return context("fp64")

    mix = rhf(m, ctx=ctx, H=H, S=S, prec="int8", e_conv=1e-8, d_conv=1e-6, switch_ctx=to_fp64)
    say("mixed RI-RHF", mix.E, mix.builds, "dE vs FP64", mix.E - ref.E, round(time.perf_counter() - t0, 1), "s")
    assert mix.converged and mix.builds.get("int8", 0) > 0 and mix.builds.get("fp64", 0) > 0
    assert abs(mix.E - ref.E) < 1e-8
