"""The ragged batch as a work-list plan (nopt_ragged_create / _launch): the
32 water clusters of the CASSCF-style batch (seed 0, two densities each)
plus Cartesian copies of the clusters of at most 3 waters, every phase of all
39 systems in one launch. The plan runs the same kernel bodies with the same
arguments as the builds launched system by system, so its outputs equal
theirs bitwise; captured in a HIP graph, every replay equals them too.
Contexts stay pinned while the plan exists."""
import sys
import time

import numpy as np
import pytest

from nopt_amd.jk import DeviceArray, Graph, JKContext, RaggedPlan, synchronize
from nopt_amd.workloads import ragged_clusters
from nopt_amd import _lib as L

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_plan_matches_system_by_system_builds():
    systems = []
    for n, m, dens in ragged_clusters(32, 16, 0):
        systems.append((n, m, dens))
    for n, m, dens in ragged_clusters(32, 16, 0, pure=False):
        if n <= 3:
            systems.append((n, m, dens))
    rows = []
    for n, m, dens in systems:
        ctx = JKContext(m.nbf, m.naux)
        ctx.set_B(m.eri3_packed(), m.metric())
        dev = [DeviceArray.from_host(a) if a.size else DeviceArray((1,)) for a in dens]
        out = [DeviceArray((m.nbf, m.nbf)) for _ in range(4)]
        rows.append((n, m, dens, ctx, dev, out))
    stream = rows[0][3].stream()
    items = [(ctx, [(dev[0].ptr, dev[1].ptr, dens[1].shape[0], out[0].ptr, out[1].ptr),
                    (dev[2].ptr, dev[3].ptr, dens[3].shape[0], out[2].ptr, out[3].ptr)])
             for n, m, dens, ctx, dev, out in rows]

    def outputs():
        return [[o.to_host() for o in r[5]] for r in rows]

    def clear():
        for r in rows:
            for o in r[5]:
                o.copy_from(np.zeros(o.shape))

    for ctx, it in items:  # system by system (sizes every buffer)
        ctx.build_batch_device(it, "int8", stream)
    synchronize()
    ref = outputs()
    clear()
    with RaggedPlan(items) as plan:
        say("plan:", len(rows), "systems,", plan.launches, "launches")
        assert plan.launches <= 16
        plan.launch(stream)
        synchronize()
        worst = max(float(np.abs(a - b).max()) for r, e in zip(outputs(), ref) for a, b in zip(r, e))
        say("plan vs system by system: max|diff|", worst)
        assert worst == 0.0
        # pinned: a build that would grow a buffer is refused
        big = rows[0]
        with pytest.raises(RuntimeError):
            C = np.eye(big[1].nbf)
            big[3].build(C, C, big[1].nbf, prec="int8")
        with Graph.capture(stream) as cap:
            plan.launch(stream)
        g = cap.graph
        for rep in range(3):
            clear()
            g.launch(stream)
            synchronize()
            worst = max(float(np.abs(a - b).max()) for r, e in zip(outputs(), ref) for a, b in zip(r, e))
            say("plan graph replay", rep, "max|diff|", worst)
            assert worst == 0.0
            rows[0][3].build(rows[0][2][0], rows[0][2][1], rows[0][2][1].shape[0], prec="int8")
        t0 = time.perf_counter()
        for _ in range(5):
            g.launch(stream)
        synchronize()
        say("plan graph: %.2f ms per batch of %d builds" % (1e3 * (time.perf_counter() - t0) / 5, 2 * len(rows)))
        g.close()
    for r in rows:
        for a in r[4] + r[5]:
            a.free()
        r[3].close()


def test_plan_rejects_bad_batches():
    """FP64-only contexts, a context twice, and orbital counts beyond the
    fused exchange split are refused at create, before any work."""
    from nopt_amd.molecule import Molecule
    from nopt_amd import geometry as G
    m = Molecule.from_angstrom(*G.water(), basis="cc-pvdz", aux="cc-pvdz-jkfit")
    D = DeviceArray((m.nbf, m.nbf))
    C = DeviceArray((m.nbf, m.nbf))
    J, K = DeviceArray((m.nbf, m.nbf)), DeviceArray((m.nbf, m.nbf))
    with JKContext(m.nbf, m.naux) as ctx, JKContext(m.nbf, m.naux, keep="fp64") as c64:
        ctx.set_B(m.eri3_packed(), m.metric())
        c64.set_B(m.eri3_packed(), m.metric())
        it = [(D.ptr, C.ptr, 5, J.ptr, K.ptr)]
        for bad in ([(c64, it)], [(ctx, it), (ctx, it)]):
            with pytest.raises(RuntimeError):
                RaggedPlan(bad)
        # four densities of 5 orbitals: 4 x 16 columns, one group (within the split)
        with RaggedPlan([(ctx, it * 4)]) as plan:
            assert plan.launches > 0
    for a in (D, C, J, K):
        a.free()
