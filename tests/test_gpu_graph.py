"""Builds of several contexts captured into one HIP graph and replayed: every
replay equals the eager build of the current buffer contents, with eager
work of the program between replays."""
import sys

import numpy as np
import pytest

from nopt_amd.molecule import Molecule
from nopt_amd.jk import JKContext, DeviceArray, Graph, synchronize
from nopt_amd import geometry as G
from systems import SYSTEMS

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def system(sym, xyz, aux, seed):
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux=aux)
    ctx = JKContext(m.nbf, m.naux)
    ctx.set_B(m.eri3_packed(), m.metric())
    rng = np.random.default_rng(seed)
    C, _ = np.linalg.qr(rng.standard_normal((m.nbf, m.nbf)))
    Co = np.ascontiguousarray(C[:, :m.nocc].T) * 0.3
    D = Co.T @ Co
    bufs = dict(D=DeviceArray.from_host(D), C=DeviceArray.from_host(Co), J=DeviceArray((m.nbf, m.nbf)),
                K=DeviceArray((m.nbf, m.nbf)))
    return m, ctx, D, Co, bufs


@pytest.mark.parametrize("prec", ["fp64", "int8"])
def test_graph_of_two_systems_matches_eager(prec):
    sysA = system(*G.water(), "cc-pvdz-jkfit", 1)
    sysB = system(*SYSTEMS["C2H4"], "cc-pvdz-ri", 2)
    s = sysA[1].stream()
    systems = (sysA, sysB)

    def enqueue(sys_):
        m, ctx, D, Co, b = sys_
        ctx.build_device(b["D"].ptr, b["C"].ptr, m.nocc, b["J"].ptr, b["K"].ptr, prec, stream=s)

    # the FP64 build run between replays below, first: a captured context's
    # buffers cannot grow while the graph exists
    sysB[1].build(sysB[2], sysB[3], sysB[0].nocc, prec="fp64")
    for x in systems:  # eager builds (also size every buffer before capture)
        enqueue(x)
    synchronize()
    eager = [(x[4]["J"].to_host(), x[4]["K"].to_host()) for x in systems]
    with Graph.capture(s) as cap:
        for x in systems:
            enqueue(x)
    graph = cap.graph
    zero = [np.zeros((x[0].nbf, x[0].nbf)) for x in systems]
    for rep in range(3):
        scale = 1.0 if rep != 1 else 0.5  # new input contents before the second replay
        for x, z in zip(systems, zero):
            m, ctx, D, Co, b = x
            b["D"].copy_from(scale * D)
            b["C"].copy_from(np.sqrt(scale) * Co)
            b["J"].copy_from(z)
            b["K"].copy_from(z)
        graph.launch(s)
        synchronize()
        for i, x in enumerate(systems):
            J, K = x[4]["J"].to_host(), x[4]["K"].to_host()
            dj = np.abs(J - scale * eager[i][0]).max()
            dk = np.abs(K - scale * eager[i][1]).max()
            say(prec, "replay", rep, "system", i, "max|dJ|", dj, "max|dK|", dk)
            tol = 1e-12 if (scale == 1.0 or prec == "fp64") else 1e-8
            assert dj <= tol * max(1.0, np.abs(eager[i][0]).max()) and dk <= tol * max(1.0, np.abs(eager[i][1]).max())
        # eager work of the program between replays
        m, ctx, D, Co, b = sysB
        ctx.build(D, Co, m.nocc, prec="fp64")
    graph.close()
    for x in systems:
        x[1].close()
