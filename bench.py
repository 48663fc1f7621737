#!/usr/bin/env python3
"""RI-JK Fock builds per second on one MI355X.

Workload: all-trans C20H42, cc-pVDZ (490 spherical functions) / cc-pVDZ-JKFIT
(2366), 81 doubly occupied orbitals. One step is one RI-J/K build (J from the
density, K from the occupied orbital factor) of the converged RI-RHF density,
with B = L^-1 (P|mn) resident in HBM and the inputs already on the device.

Every precision mode is timed (J is FP64 in all of them): fp64, int8 (int8
digit splits with power-of-two row scales, exact int32 MFMA sums), fp32
(FP32 MFMA products, FP64 accumulation), bf16x6 (three-piece bf16 split of
FP32 operands, six MFMA products, FP64 accumulation) and bf16x3. The headline
value is the fastest mode whose energy error against the FP64 build on the
same density stays below 1 micro-Hartree (--prec auto), next to the FP64
build's own rate. Accuracy is also reported on ten perturbed densities
D + 1e-3 sym(R) (seed 0), whose K comes from a signed eigen-factor.

The CPU baseline times the oracle's restatement of NOPT's
DM_to_F_transform_RI (Cartesian AOs, raw (mn|P), V^-1/2 applied per build,
OpenMP row-split MKL DGEMMs) on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# the form of B each mode reads (all are formed once per geometry from the
# same FP64 B = L^-1 (P|mn))
B_FORM = {"int8": "int8 digit planes (4 planes of the full rows for the half-transform, 5-byte packed "
                  "lower triangle for gamma and J)",
          "fp32": "FP64 rows (rounded to FP32 on the fly)", "bf16x6": "FP64 rows (split to bf16 on the fly)",
          "bf16x3": "FP64 rows (split to bf16 on the fly)", "fp64": "FP64 rows"}
B_FORM_COMPACT = "a 5-byte packed lower triangle (28-bit integer + last digit), read by every kernel"
PEAK = {"hbm_GBps": 8000.0, "fp64_TF": 78.6, "fp32_TF": 157.3, "bf16_TF": 2500.0, "int8_TOPS": 5000.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def energy_terms(D, J, K):
    """Two-electron part of the RHF energy, sum D.(2J - K)."""
    return float(np.sum(D * (2.0 * J - K)))


def time_builds(ctx, dD, dC, nocc, dJ, dK, prec, steps, warmup, synchronize, timing=False):
    """Builds/s of `steps` builds after `warmup` (device synchronised on both
    sides). With timing, a second pass of `steps` builds records HIP events
    around each phase; phases then run one after the other on one stream,
    while the timed pass lets the Coulomb pair overlap the exchange."""
    for _ in range(warmup):
        ctx.build_device(dD.ptr, dC.ptr, nocc, dJ.ptr, dK.ptr, prec)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.build_device(dD.ptr, dC.ptr, nocc, dJ.ptr, dK.ptr, prec)
    synchronize()
    dt = time.perf_counter() - t0
    phases = None
    if timing:
        ctx.set_timing(True)
        for _ in range(steps):
            ctx.build_device(dD.ptr, dC.ptr, nocc, dJ.ptr, dK.ptr, prec)
        synchronize()
        ph, nb = ctx.read_timing()
        ctx.set_timing(False)
        phases = {k: v / max(nb, 1) for k, v in ph.items()}
    return steps / dt, 1000.0 * dt / steps, phases


# kernels of each timed phase, per mode
KERNELS = {
    "half_transform": {"fp64": "k_half<double, 6>", "fp32": "k_half<float, 6>", "bf16x6": "k_half_bf16<3, 6>",
                       "bf16x3": "k_half_bf16<2, 6>", "int8": "k_half_i8<6, 2, 4, 2, 32, 4, false>"},
    "gamma": {"fp64": "k_gamma (fused partials)", "fp32": "k_gamma (fused partials)",
              "bf16x6": "k_gamma (fused partials)", "bf16x3": "k_gamma (fused partials)",
              "int8": "k_gamma_slots<1, 2, true>"},
    "coulomb_j": {"fp64": "k_jpart<1>", "fp32": "k_jpart<1>", "bf16x6": "k_jpart<1>", "bf16x3": "k_jpart<1>",
                  "int8": "k_jpart_slots<1, 2, true>"},
    "exchange": {"fp64": "k_syrk<double>", "fp32": "k_syrk<float>", "bf16x6": "k_syrk_bf16<3>",
                 "bf16x3": "k_syrk_bf16<2>", "int8": "k_syrk_i8<4, true>"},
}


def roofline(phase, ms, nbf, naux, nocc, prec, half_packed=False, shares=None):
    """Algorithmic work of the phase's kernel per launch, over its HIP-event time.
    Flops are the FP64 contraction's (F_half = 2 Q nbf^2 o, F_K = Q o nbf(nbf+1),
    one triangle), whatever arithmetic carries them, against the peak of the
    arithmetic used; the int8 modes' digit-pair products (13 per product) are
    reported beside as the emulation's own rate. Bytes count each operand
    element once at its stored width (no padding): B in the form the kernel
    reads (full rows or packed lower triangle), C, and X written. The int8
    kernels leave out B's blocks of exact zeros: their B bytes and the
    half-transform's products count the share they read (`shares`, from
    nopt_jk_zero_blocks: slabs for the half-transform, segments for the
    Coulomb pair), and `dense_equivalent` gives the rate counted over all of B."""
    sh = shares or {}
    s_slab, s_seg = sh.get("slabs", 1.0), sh.get("segments", 1.0)
    t = ms * 1e-3
    npair = nbf * (nbf + 1) / 2

    def hbm(byts):
        a = byts / t / 1e9
        return dict(bound="hbm", achieved=a, peak=PEAK["hbm_GBps"], unit="GB/s", frac=a / PEAK["hbm_GBps"],
                    traffic=None, kernel=KERNELS[phase][prec], algorithmic_bytes=byts)

    def mfma(ops, peak, unit="TFLOP/s"):
        a = ops / t / 1e12
        return dict(bound="mfma", achieved=a, peak=peak, unit=unit, frac=a / peak, traffic=None,
                    kernel=KERNELS[phase][prec], algorithmic_flops=ops)

    f_half = 2.0 * naux * nbf * nbf * nocc
    f_k = float(naux * nocc * nbf * (nbf + 1))
    if phase == "half_transform":
        if prec == "fp64":
            return mfma(f_half, PEAK["fp64_TF"])
        if prec == "int8":  # 4 B digit planes in (the slabs read), 4 X digit planes out
            bb = 4.0 * naux * (npair if half_packed else nbf * nbf)
            r = hbm(s_slab * bb + 4.0 * naux * nbf * nocc)
            r["b_form"] = ("packed lower triangle (28-bit integers, unpacked to 4 digit planes in LDS)"
                           if half_packed else "4 int8 digit planes, full rows")
            r["share_of_B_read"] = s_slab
            r["dense_equivalent"] = dict(bytes=bb + 4.0 * naux * nbf * nocc,
                                         achieved_GBps=(bb + 4.0 * naux * nbf * nocc) / t / 1e9)
            fl = s_slab * f_half
            r["mfma_view"] = dict(algorithmic_flops=fl, achieved_TOPs=fl / t / 1e12,
                                  frac_of_int8_peak=fl / t / 1e12 / PEAK["int8_TOPS"],
                                  emulation_frac=13 * fl / t / 1e12 / PEAK["int8_TOPS"],
                                  dense_equivalent_TOPs=f_half / t / 1e12)
            return r
        return hbm(8.0 * naux * nbf * nbf + 4.0 * naux * nbf * nocc)  # FP64 B in, FP32 X out
    if phase == "exchange":
        if prec == "int8":
            r = mfma(f_k, PEAK["int8_TOPS"], "TOP/s")
            # the 13 digit-pair products per product on the int8 matrix cores,
            # and the rate as the FP64 work it replaces
            r["emulation"] = dict(int8_ops=13 * f_k, achieved_TOPs=13 * f_k / t / 1e12,
                                  frac=13 * f_k / t / 1e12 / PEAK["int8_TOPS"])
            r["fp64_equivalent"] = dict(achieved_TFLOPs=f_k / t / 1e12,
                                        frac_of_fp64_peak=f_k / t / 1e12 / PEAK["fp64_TF"])
            return r
        pk = {"fp64": PEAK["fp64_TF"], "fp32": PEAK["fp32_TF"], "bf16x6": PEAK["bf16_TF"] / 6,
              "bf16x3": PEAK["bf16_TF"] / 3}[prec]
        return mfma(f_k, pk)
    if phase in ("gamma", "coulomb_j") and prec == "int8":
        # lower triangle: int32 high plane + signed-byte low digit, the live segments
        r = hbm(5.0 * naux * npair * s_seg)
        r["share_of_B_read"] = s_seg
        r["dense_equivalent"] = dict(bytes=5.0 * naux * npair, achieved_GBps=5.0 * naux * npair / t / 1e9)
        return r
    if phase == "gamma":
        return None  # partials ride on the half-transform
    if phase == "coulomb_j":
        return hbm(8.0 * naux * npair)
    return None


def perturbed_accuracy(ctx, D, modes, seed=0, count=10, eps=1e-3):
    """J/K of D_k = D + eps sym(R_k), R_k ~ N(0,1) from one seeded generator.
    D_k is indefinite, so K comes from its signed eigen-factor in one batched
    build, K = K(X+) - K(X-); J from D_k. Max |dJ|, |dK| and |dE| of each mode
    against the FP64 build of the same D_k."""
    rng = np.random.default_rng(seed)
    n = D.shape[0]
    out = {p: dict(max_abs_dJ=0.0, max_abs_dK=0.0, max_abs_dE=0.0) for p in modes if p != "fp64"}
    for _ in range(count):
        R = rng.standard_normal((n, n))
        Dk = D + eps * 0.5 * (R + R.T)
        s, X = np.linalg.eigh(Dk)
        Xp = np.ascontiguousarray((X[:, s > 0] * np.sqrt(s[s > 0])).T)
        Xm = np.ascontiguousarray((X[:, s < 0] * np.sqrt(-s[s < 0])).T)
        res = {}
        for p in modes:
            (J, Kp), (_, Km) = ctx.build_batch([(Dk, Xp, Xp.shape[0]), (None, Xm, Xm.shape[0])], prec=p)
            res[p] = (J, Kp - Km)
        J0, K0 = res["fp64"]
        E0 = energy_terms(Dk, J0, K0)
        for p in out:
            J, K = res[p]
            o = out[p]
            o["max_abs_dJ"] = max(o["max_abs_dJ"], float(np.abs(J - J0).max()))
            o["max_abs_dK"] = max(o["max_abs_dK"], float(np.abs(K - K0).max()))
            o["max_abs_dE"] = max(o["max_abs_dE"], abs(energy_terms(Dk, J, K) - E0))
    return out


def ragged_batch(prec, steps, warmup, count=32, nmax=16, seed=0):
    """The CASSCF-style ragged batch: `count` water clusters (H2O)_n, n ~
    U{1..nmax} (seed), cc-pVDZ/cc-pVDZ-JKFIT, each with an inactive density
    (all but 5 doubly occupied core-guess orbitals) and a total density with
    an active space of 10 orbitals and a random PSD 1-RDM of trace 10
    (factor [sqrt(2) C_core; NO sqrt(n/2)], src/molecule.cpp:2959-2977). All
    2*count builds are captured in one HIP graph from one
    nopt_jk_build_ragged_device call; builds/s of the replayed graph and of the
    builds launched one by one, with the split counts of each system set by
    the autotuner (the replay with the counts picked at context creation is
    reported beside it)."""
    from nopt_amd.jk import JKContext, DeviceArray, Graph, RaggedPlan, synchronize, build_ragged_device
    from nopt_amd.workloads import ragged_clusters
    t0 = time.perf_counter()
    systems, sizes = [], []
    for n, m, (Dc, Cc, Dt, Ct) in ragged_clusters(count, nmax, seed):
        ctx = JKContext(m.nbf, m.naux)
        ctx.set_B(m.eri3_packed(), m.metric())
        d = dict(m=m, ctx=ctx, ncore=Cc.shape[0], nt=Ct.shape[0], host=(Dc, Cc, Dt, Ct))
        d["dev"] = [DeviceArray.from_host(a) if a.size else DeviceArray((1,)) for a in (Dc, Cc, Dt, Ct)]
        d["out"] = [DeviceArray((m.nbf, m.nbf)) for _ in range(4)]
        systems.append(d)
        sizes.append(n)
    setup = time.perf_counter() - t0
    stream = systems[0]["ctx"].stream()

    def enqueue(p):
        for d in systems:
            dDc, dCc, dDt, dCt = d["dev"]
            Jc, Kc, Jt, Kt = d["out"]
            d["ctx"].build_batch_device([(dDc.ptr, dCc.ptr, d["ncore"], Jc.ptr, Kc.ptr),
                                         (dDt.ptr, dCt.ptr, d["nt"], Jt.ptr, Kt.ptr)], p, stream)

    def ragged(p):
        # every system in one call
        build_ragged_device([(d["ctx"], [(d["dev"][0].ptr, d["dev"][1].ptr, d["ncore"], d["out"][0].ptr,
                                          d["out"][1].ptr),
                                         (d["dev"][2].ptr, d["dev"][3].ptr, d["nt"], d["out"][2].ptr,
                                          d["out"][3].ptr)]) for d in systems], p, stream)

    def graph_ms(p, fn=None):
        fn = fn or ragged
        fn(p)  # sizes every buffer before the capture
        synchronize()
        with Graph.capture(stream) as cap:
            fn(p)
        g = cap.graph
        g.launch(stream)
        synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            g.launch(stream)
        synchronize()
        g.close()
        return 1e3 * (time.perf_counter() - t1) / steps

    # the split counts of every system chosen by the autotuner (device builds
    # of its total density), after a replay with the counts create picked
    untuned = {p: graph_ms(p) for p in ("fp64", prec)}
    t0 = time.perf_counter()
    for d in systems:
        dDc, dCc, dDt, dCt = d["dev"]
        for p in ("fp64", prec):
            d["ctx"].autotune(dDt.ptr, dCt.ptr, d["nt"], prec=p, reps=2)
    tune_s = time.perf_counter() - t0
    res = {}
    for p in ("fp64", prec):
        for _ in range(warmup):
            enqueue(p)
        synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            enqueue(p)
        synchronize()
        eager = (time.perf_counter() - t1) / steps
        ref = [[o.to_host() for o in d["out"]] for d in systems]
        with Graph.capture(stream) as cap:
            ragged(p)
        g = cap.graph
        g.launch(stream)
        synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            g.launch(stream)
        synchronize()
        graph = (time.perf_counter() - t1) / steps
        same = max(float(np.abs(o.to_host() - r).max()) for d, rr in zip(systems, ref) for o, r in zip(d["out"], rr))
        g.close()
        res[p] = dict(eager_ms=1e3 * eager, graph_ms=1e3 * graph, builds_per_s_graph=2 * count / graph,
                      builds_per_s_eager=2 * count / eager, graph_vs_eager_max_abs_diff=same,
                      graph_ms_default_splits=untuned[p], outputs=ref)
        if p == "int8":
            # the work-list plan: every phase of all systems in one launch
            for d in systems:
                for o in d["out"]:
                    o.copy_from(np.zeros(o.shape))
            with RaggedPlan([(d["ctx"], [(d["dev"][0].ptr, d["dev"][1].ptr, d["ncore"], d["out"][0].ptr,
                                          d["out"][1].ptr),
                                         (d["dev"][2].ptr, d["dev"][3].ptr, d["nt"], d["out"][2].ptr,
                                          d["out"][3].ptr)]) for d in systems]) as plan:
                plan.launch(stream)
                synchronize()
                same_plan = max(float(np.abs(o.to_host() - r).max())
                                for d, rr in zip(systems, ref) for o, r in zip(d["out"], rr))
                with Graph.capture(stream) as cap:
                    plan.launch(stream)
                g = cap.graph
                g.launch(stream)
                synchronize()
                t1 = time.perf_counter()
                for _ in range(steps):
                    g.launch(stream)
                synchronize()
                tp = (time.perf_counter() - t1) / steps
                g.close()
                res[p].update(plan_graph_ms=1e3 * tp, builds_per_s_plan=2 * count / tp, plan_launches=plan.launches,
                              plan_vs_eager_max_abs_diff=same_plan)
    # accuracy of the mixed mode against FP64 over the whole batch, as the
    # two-electron energy of each density: sum Dc.(2J - K) for the inactive
    # density (unit occupations, closed shell) and 1/2 sum Dt.J - 1/4 sum Dt.K
    # for the spin-summed total density (occupations 0..2), which is a
    # quarter of sum Dt.(2J - K)
    dE = dK = 0.0
    worst = {}
    for d, n, r64, rp in zip(systems, sizes, res["fp64"]["outputs"], res[prec]["outputs"]):
        Dc, Cc, Dt, Ct = d["host"]
        for kind, w, D, J64, K64, J, K in (("inactive", 1.0, Dc, r64[0], r64[1], rp[0], rp[1]),
                                           ("total", 0.25, Dt, r64[2], r64[3], rp[2], rp[3])):
            e64 = w * energy_terms(D, J64, K64)
            de = abs(w * energy_terms(D, J, K) - e64)
            if de > worst.get(kind, {}).get("dE", -1.0):
                worst[kind] = dict(dE=de, E2=e64, waters=int(n))
            dE = max(dE, de)
            dK = max(dK, float(np.abs(K - K64).max()))
    for p in res:
        res[p].pop("outputs")
    for d in systems:
        for a in d["dev"] + d["out"]:
            a.free()
        d["ctx"].close()
    return dict(clusters=count, waters=int(sum(sizes)), sizes=sizes, builds_per_graph=2 * count, autotune_s=round(tune_s, 1),
                nbf_range=[min(d["m"].nbf for d in systems), max(d["m"].nbf for d in systems)],
                setup_s=round(setup, 1), modes={p: {k: round(v, 4) for k, v in r.items()} for p, r in res.items()},
                max_abs_dE_vs_fp64=dE, max_abs_dK_vs_fp64=dK, worst_dE_by_density=worst, precision=prec)


def large_system(n_water, seed, steps, threads):
    """The large single system resident in HBM: a water cluster (H2O)_n
    (seed, 1 g/cm^3, O-O >= 2.7 A) in cc-pVDZ / cc-pVDZ-JKFIT. B does not fit
    twice in HBM (FP64 rows and digit planes), so each context keeps one form:
    an FP64 context runs the reference RI-RHF to 1e-8 Eh from the core guess;
    an integer context runs the same SCF with integer builds until the DIIS
    error falls below 1e-5, then hands over to an FP64 context for the
    end-game. Reports the energy difference of the two runs, the J/K/energy
    differences of single builds on the converged density, and the builds/s
    of both modes (device buffers, `steps` builds after one warm-up)."""
    import contextlib
    from nopt_amd import geometry as G
    from nopt_amd.molecule import Molecule
    from nopt_amd.jk import JKContext, DeviceArray, synchronize
    from nopt_amd.scf import rhf
    from nopt_amd import _lib as L
    L.lib().nopt_set_host_threads(threads)
    sym, xyz = G.water_cluster(n_water, seed)
    mol = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit")
    nbf, naux, nocc = mol.nbf, mol.naux, mol.nocc
    log(f"[large] (H2O){n_water} seed {seed}: nbf {nbf} naux {naux} nocc {nocc}; host integrals ...")
    t0 = time.perf_counter()
    S, T, V = mol.one_electron()
    H = T + V
    metric = mol.metric()
    eri3 = mol.eri3_packed()
    t_ints = time.perf_counter() - t0
    log(f"[large] host integrals {t_ints:.1f} s, packed (P|mn) {eri3.nbytes / 1e9:.1f} GB")
    timings = {}

    def context(keep):
        t1 = time.perf_counter()
        c = JKContext(nbf, naux, keep=keep)
        c.set_B(eri3, metric)
        dt = time.perf_counter() - t1
        timings.setdefault(f"set_B_{keep}_s", []).append(round(dt, 1))
        log(f"[large] {keep} context: B formed in {dt:.1f} s, {c.device_bytes() / 1e9:.1f} GB on the device")
        return c

    def rate(c, D, Co, prec):
        dD, dC = DeviceArray.from_host(D), DeviceArray.from_host(Co)
        dJ, dK = DeviceArray((nbf, nbf)), DeviceArray((nbf, nbf))
        r, _, ph = time_builds(c, dD, dC, nocc, dJ, dK, prec, steps, 1, synchronize, timing=True)
        timings[f"phase_ms_{prec}"] = {k: round(v, 3) for k, v in ph.items()}
        timings[f"device_bytes_{prec}_context"] = c.device_bytes()
        out = (dJ.to_host(), dK.to_host())
        for a in (dD, dC, dJ, dK):
            a.free()
        return r, out

    ctx = context("fp64")
    t0 = time.perf_counter()
    with contextlib.redirect_stdout(sys.stderr):
        ref = rhf(mol, ctx=ctx, H=H, S=S, prec="fp64", e_conv=1e-8, d_conv=1e-6, verbose=True)
    t_ref = time.perf_counter() - t0
    log(f"[large] FP64 RI-RHF E = {ref.E:.10f} in {ref.iters + 1} iterations, {t_ref:.1f} s")
    Co = np.ascontiguousarray(ref.C[:, :nocc].T)
    D = Co.T @ Co
    rate64, (J64, K64) = rate(ctx, D, Co, "fp64")
    log(f"[large] fp64: {rate64:.3f} builds/s; phases {timings['phase_ms_fp64']}")
    ctx.close()
    ctx = context("int8")
    timings["zero_blocks"] = ctx.zero_blocks()
    log(f"[large] shares read after the exact-zero blocks are left out: {timings['zero_blocks']}")
    rate8, (J8, K8) = rate(ctx, D, Co, "int8")
    dE_build = energy_terms(D, J8, K8) - energy_terms(D, J64, K64)
    log(f"[large] int8: {rate8:.3f} builds/s; phases {timings['phase_ms_int8']}, "
        f"{timings['device_bytes_int8_context'] / 1e9:.1f} GB on the device; on the converged density dE {dE_build:.3e} "
        f"max|dJ| {np.abs(J8 - J64).max():.3e} max|dK| {np.abs(K8 - K64).max():.3e}")

    def to_fp64(c):
        c.close()
        return context("fp64")

    t0 = time.perf_counter()
    with contextlib.redirect_stdout(sys.stderr):
        mix = rhf(mol, ctx=ctx, H=H, S=S, prec="int8", e_conv=1e-8, d_conv=1e-6, verbose=True, switch_ctx=to_fp64)
    t_mix = time.perf_counter() - t0
    log(f"[large] mixed RI-RHF E = {mix.E:.10f} ({mix.builds}), dE vs FP64 run {mix.E - ref.E:.3e}, {t_mix:.1f} s")
    return {
        "metric": f"RI-RHF (H2O){n_water} cc-pVDZ/cc-pVDZ-JKFIT, integer builds with FP64 end-game vs FP64",
        "value": round(rate8, 4), "unit": "builds/s", "n_gpus": 1, "higher_is_better": True, "dtype": "int8",
        "data": f"water cluster seed {seed}, random orientations, 1 g/cm^3, O-O >= 2.7 A",
        "config": {"workload": f"(H2O){n_water}", "nbf": nbf, "naux": naux, "nocc": nocc,
                   "fp64_builds_per_s": round(rate64, 4), "int8_vs_fp64_build_speedup": round(rate8 / rate64, 3),
                   "E_fp64": ref.E, "E_mixed": mix.E, "dE_scf": mix.E - ref.E,
                   "scf_fp64": {"iterations": ref.iters + 1, "builds": ref.builds, "seconds": round(t_ref, 1)},
                   "scf_mixed": {"iterations": mix.iters + 1, "builds": mix.builds, "switch_iter": mix.switch_iter,
                                 "seconds": round(t_mix, 1)},
                   "build_on_converged_density": {"dE": dE_build, "max_abs_dJ": float(np.abs(J8 - J64).max()),
                                                  "max_abs_dK": float(np.abs(K8 - K64).max())},
                   "host_integrals_s": round(t_ints, 1), **timings,
                   "B_bytes": {"fp64_rows": 8 * naux * nbf * (((nbf + 1 + 15) // 16) * 16),
                               "int8_packed_triangle": 5 * naux * 16 * seg_count(nbf)}},
    }


def seg_count(nbf):
    """16-column segments of the packed lower triangle of one B_Q (row m holds
    m // 16 + 1 of them)."""
    return sum(m // 16 + 1 for m in range(nbf))


def host_cores():
    """Physical cores this process may run on: the distinct (package, core)
    pairs of the CPUs in its affinity mask (sysfs topology), capped by the
    host's CPU share for this job (OMP_NUM_THREADS when the environment sets
    it). Returns (cores, description)."""
    cpus = sorted(os.sched_getaffinity(0))
    phys = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            phys.add((open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip()))
        except OSError:
            phys.add(("?", str(c)))
    n = len(phys)
    share = os.environ.get("OMP_NUM_THREADS", "")
    cap = int(share) if share.isdigit() and int(share) > 0 else None
    cores = min(n, cap) if cap else n
    desc = (f"{n} physical cores ({len(cpus)} logical CPUs) in the affinity mask"
            + (f", CPU share OMP_NUM_THREADS={cap}" if cap else "") + f": {cores} threads")
    return cores, desc


def converged_cartesian(sym, xyz_ang):
    """The converged FP64 RI-RHF orbitals of the workload in NOPT's Cartesian
    representation (the basis of the CPU path), from GPU builds: (DM, MO) with
    MO rows the occupied orbitals."""
    from nopt_amd.molecule import Molecule
    from nopt_amd.jk import JKContext
    from nopt_amd.scf import rhf
    mc = Molecule.from_angstrom(sym, xyz_ang, pure=False)
    with JKContext(mc.nbf, mc.naux) as ctx:
        ctx.set_B(mc.eri3_packed(), mc.metric())
        r = rhf(mc, ctx=ctx, prec="fp64", e_conv=1e-9, d_conv=1e-7)
    MO = np.ascontiguousarray(r.C[:, :mc.nocc].T)
    return MO.T @ MO, MO, r.E


def cpu_baseline(sym, xyz_ang, threads, DM, MO, nbuilds=5, label="C20H42", cores_desc=""):
    """NOPT's CPU path (oracle restatement) on a bounded sample: one warm-up
    and `nbuilds` timed builds of the full C20H42 Cartesian workload (V^-1/2
    applied per build, as DM_to_F_transform_RI does) at the converged density
    and occupied orbitals (DM, MO), then the same for the "B-prepared"
    variant, whose stored tensor already holds the metric."""
    os.environ["ORACLE_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from nopt_amd.molecule import Molecule
    from nopt_amd import _lib as L
    L.lib().nopt_set_host_threads(threads)
    t0 = time.perf_counter()
    mc = Molecule.from_angstrom(sym, xyz_ang, pure=False)  # NOPT reads the orbital basis Cartesian
    RI_AA = mc.eri3_nopt()
    Vm05, _, _ = O.s05(mc.metric())
    O.lib().oracle_set_threads(threads)
    assert MO.shape == (mc.nocc, mc.nbf) and DM.shape == (mc.nbf, mc.nbf)
    setup = time.perf_counter() - t0

    def median_s(jk):
        jk(DM, MO, mc.nocc)
        ts = []
        for _ in range(nbuilds):
            t1 = time.perf_counter()
            jk(DM, MO, mc.nocc)
            ts.append(time.perf_counter() - t1)
        return float(np.median(ts))

    jk = O.NoptJK(RI_AA, Vm05, mc.nocc)
    med = median_s(jk)
    del jk
    t1 = time.perf_counter()
    jkp = O.NoptJKPrepared(RI_AA, Vm05, mc.nocc)
    prep_s = time.perf_counter() - t1
    del RI_AA
    medp = median_s(jkp)
    del jkp
    return dict(value=1.0 / med, unit="builds/s", cores=threads, kind="port",
                sample=f"{nbuilds} timed builds after 1 warm-up of the oracle's C restatement of NOPT's "
                       f"DM_to_F_transform_RI (MKL DGEMM, OpenMP row split as nopt_par_dgemm) on {label}: "
                       f"Cartesian n_ao {mc.nbf}, naux {mc.naux}, n_c {mc.nocc}, converged FP64 RI-RHF "
                       f"density and orbitals; median {med:.3f} s/build",
                cores_detail=cores_desc, blas="MKL (libmkl_rt)",
                b_prepared=dict(value=1.0 / medp, unit="builds/s", median_s=round(medp, 4), prepare_s=round(prep_s, 1),
                                sample=f"{nbuilds} timed builds after 1 warm-up, same workload with RI_AA.V^-1/2 "
                                       f"formed once (no metric products per build)"),
                setup_s=round(setup, 1))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--prec", default="auto", choices=["auto", "int8", "fp32", "bf16x6", "bf16x3", "fp64"],
                    help="headline mode; auto = fastest mode with |dE| < 1 uEh against FP64")
    ap.add_argument("--alkane", type=int, default=20, help="carbon count of the all-trans alkane")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tuning", default="", metavar="KEY=N,...",
                    help="split counts for the modes other than fp64 instead of the autotuner, e.g. "
                         "nsk8=8,nsg=64,nsj8=8,cplace=0 (keys of JKContext.TUNING_KEYS)")
    ap.add_argument("--modes", default="fp64,int8",
                    help="comma-separated modes to time (fp64 is always built as the reference); the FP32 and "
                         "bf16 modes (fp32, bf16x6, bf16x3) are slower and less accurate than int8 at C20H42 and "
                         "are timed only when named")
    ap.add_argument("--no-perturbed", action="store_true", help="skip the perturbed-density accuracy check")
    ap.add_argument("--no-ragged", action="store_true", help="skip the ragged water-cluster batch")
    ap.add_argument("--compact", action="store_true",
                    help="integer form of B as the packed triangle alone (the large-system form)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (HBM counter calibration) and exit")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (default: the physical cores of this job's CPU share)")
    ap.add_argument("--large", type=int, default=0, metavar="NWATER",
                    help="instead of the headline, run the large-system RI-RHF on (H2O)_NWATER (seed 1) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only (--gpus 1); sharding the build over the auxiliary index "
              "across GPUs is designed but not implemented, so no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        from nopt_amd.jk import DeviceArray, synchronize
        n = args.copy_probe * (1 << 20) // 8
        a, b = DeviceArray((n,)), DeviceArray((n,))
        from nopt_amd import _lib as L
        for _ in range(5):
            L.check(L.lib().nopt_dev_memcpy(b.ptr, a.ptr, a.nbytes, 2))
        synchronize()
        print(json.dumps({"copy_probe_bytes": a.nbytes, "copies": 5}))
        return

    cores, cores_desc = host_cores()
    if args.cpu_threads <= 0:
        args.cpu_threads = cores
    else:
        cores_desc = f"{args.cpu_threads} threads (--cpu-threads); " + cores_desc
    if args.large:
        print(json.dumps(large_system(args.large, 1, max(3, min(args.steps, 10)), args.cpu_threads)))
        return

    from nopt_amd import geometry as G
    from nopt_amd.molecule import Molecule
    from nopt_amd.jk import JKContext, DeviceArray, synchronize
    from nopt_amd.scf import rhf
    from nopt_amd import _lib as L

    L.lib().nopt_set_host_threads(args.cpu_threads)
    sym, xyz = G.alkane(args.alkane)
    t0 = time.perf_counter()
    mol = Molecule.from_angstrom(sym, xyz)
    S, T, V = mol.one_electron()
    eri3, metric = mol.eri3_packed(), mol.metric()
    t_ints = time.perf_counter() - t0
    log(f"[bench] C{args.alkane}: nbf {mol.nbf} naux {mol.naux} nocc {mol.nocc}; host integrals {t_ints:.1f} s")
    ctx = JKContext(mol.nbf, mol.naux, compact=args.compact)
    t0 = time.perf_counter()
    ctx.set_B(eri3, metric)
    t_B = time.perf_counter() - t0
    del eri3
    log(f"[bench] B formed on the device (Cholesky + triangular solve) in {t_B:.2f} s")
    zero_blocks = ctx.zero_blocks()
    log(f"[bench] shares of the integer forms read after the exact-zero blocks are left out: {zero_blocks}")
    t0 = time.perf_counter()
    res = rhf(mol, ctx=ctx, H=T + V, S=S, prec="fp64", e_conv=1e-10, d_conv=1e-8)
    t_scf = time.perf_counter() - t0
    log(f"[bench] FP64 RI-RHF E = {res.E:.10f} in {res.iters + 1} iterations, {t_scf:.1f} s")

    nocc, nbf, naux = mol.nocc, mol.nbf, mol.naux
    Co = np.ascontiguousarray(res.C[:, :nocc].T)
    D = Co.T @ Co
    dD, dC = DeviceArray.from_host(D), DeviceArray.from_host(Co)
    dJ, dK = DeviceArray((nbf, nbf)), DeviceArray((nbf, nbf))

    # accuracy of each mode against the FP64 build on the same density
    modes = tuple(["fp64"] + [p for p in args.modes.split(",") if p and p != "fp64"])
    fixed_tuning = {}
    for kv in filter(None, args.tuning.split(",")):
        k, v = kv.split("=")
        if k not in JKContext.TUNING_KEYS:
            raise SystemExit(f"--tuning: unknown key {k} (one of {', '.join(JKContext.TUNING_KEYS)})")
        fixed_tuning[k] = int(v)
    acc = {}
    ref = None
    for p in modes:
        ctx.build_device(dD.ptr, dC.ptr, nocc, dJ.ptr, dK.ptr, p)
        synchronize()
        J, K = dJ.to_host(), dK.to_host()
        if ref is None:
            ref = (J, K, energy_terms(D, J, K))
        acc[p] = dict(dE=energy_terms(D, J, K) - ref[2], max_abs_dJ=float(np.abs(J - ref[0]).max()),
                      max_abs_dK=float(np.abs(K - ref[1]).max()))
    log(f"[bench] accuracy vs FP64: {json.dumps(acc)}")
    pert = None if args.no_perturbed else perturbed_accuracy(ctx, D, modes, seed=0, count=10, eps=1e-3)
    log(f"[bench] perturbed densities (10, eps 1e-3): {json.dumps(pert)}")

    runs, tuning = {}, {}
    for p in modes:
        # split counts chosen by the autotuner on this density, then the timed builds
        # (16 builds per candidate: the counts that matter differ by a few per cent)
        if fixed_tuning and p != "fp64":
            ctx.set_tuning(**fixed_tuning)
            tuning[p] = ctx.tuning()
        else:
            tuning[p], _ = ctx.autotune(dD.ptr, dC.ptr, nocc, prec=p, reps=16)
        runs[p] = time_builds(ctx, dD, dC, nocc, dJ, dK, p, args.steps, args.warmup, synchronize, timing=True)
        log(f"[bench] {p}: {runs[p][0]:.1f} builds/s ({runs[p][1]:.3f} ms/build); phases {runs[p][2]}")
    if args.prec == "auto":
        ok = [p for p in modes if p != "fp64" and abs(acc[p]["dE"]) < 1e-6]
        prec = max(ok, key=lambda p: runs[p][0]) if ok else "fp64"
    else:
        prec = args.prec
    rate, ms, phases = runs[prec]
    rate64 = runs["fp64"][0]
    # host-buffer path (D, C in and J, K out over PCIe, synchronous): reported, never the value
    Ch = np.ascontiguousarray(Co)
    ctx.build(D, Ch, nocc, prec=prec)
    t0 = time.perf_counter()
    nh = max(5, args.steps // 5)
    for _ in range(nh):
        ctx.build(D, Ch, nocc, prec=prec)
    host_rate = nh / (time.perf_counter() - t0)
    log(f"[bench] host-buffer path ({prec}): {host_rate:.1f} builds/s")

    timed = ("half_transform", "gamma", "coulomb_j", "exchange")
    hp = args.compact
    zb = zero_blocks
    dom = max((k for k in timed if roofline(k, phases[k], nbf, naux, nocc, prec, hp, zb)), key=lambda k: phases[k])
    roof = roofline(dom, phases[dom], nbf, naux, nocc, prec, hp, zb)
    kern = {p: {k: roofline(k, runs[p][2][k], nbf, naux, nocc, p, hp, zb) for k in timed} for p in modes}
    # measured HBM bytes per dispatch (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated) for this workload
    tfile = os.path.join(ROOT, "profiles", "hbm_traffic_c20h42.json")
    if args.alkane == 20 and os.path.exists(tfile):
        meas = json.load(open(tfile))["kernels"]
        for p in modes:
            for r in kern[p].values():
                if r and r["kernel"] in meas:
                    r["traffic"] = meas[r["kernel"]]["traffic_bytes"]
        if roof["kernel"] in meas:
            roof["traffic"] = meas[roof["kernel"]]["traffic_bytes"]

    ragged = None
    if not args.no_ragged:
        ragged = ragged_batch(prec, max(3, args.steps // 5), 1)
        log(f"[bench] ragged batch ({ragged['clusters']} clusters, {ragged['builds_per_graph']} builds per graph): "
            f"{json.dumps(ragged['modes'])}; worst |dE| vs fp64 {json.dumps(ragged['worst_dE_by_density'])}")

    cpu = None
    if not args.no_cpu_baseline:
        ctx.close()
        for a in (dD, dC, dJ, dK):
            a.free()
        t0 = time.perf_counter()
        DMc, MOc, Ec = converged_cartesian(sym, xyz)
        log(f"[bench] Cartesian FP64 RI-RHF for the CPU path: E = {Ec:.10f} ({time.perf_counter() - t0:.1f} s)")
        cpu = cpu_baseline(sym, xyz, args.cpu_threads, DMc, MOc, label=f"C{args.alkane}H{2 * args.alkane + 2}",
                           cores_desc=cores_desc)
        log(f"[bench] CPU baseline {cpu['value']:.3f} builds/s on {cpu['cores']} threads, B-prepared variant "
            f"{cpu['b_prepared']['value']:.3f} builds/s ({time.perf_counter() - t0:.1f} s)")

    line = {
        "metric": "RI-JK Fock builds/s (C20H42 cc-pVDZ), mixed vs FP64 GPU at |dE|<1 uEh",
        "value": round(rate, 2),
        "unit": "builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"int8": "int8", "fp32": "f32", "bf16x6": "bf16", "bf16x3": "bf16", "fp64": "f64"}[prec],
        "data": "C20H42 geometry built in-repo; converged FP64 RI-RHF density and orbitals",
        "config": {
            "workload": f"RI-J/K build, all-trans C{args.alkane}H{2 * args.alkane + 2} cc-pVDZ/cc-pVDZ-JKFIT, "
                        f"nbf {nbf}, naux {naux}, nocc {nocc}, B resident in HBM as "
                        f"{B_FORM_COMPACT if (args.compact and prec == 'int8') else B_FORM[prec]}",
            "precision": prec,
            "arithmetic": {"int8": "int8 digit products, exact int32 class sums, fp64 combine; J from 35-bit "
                                   "integers in fp64",
                           "fp32": "f32 products, f32 chains of 16, fp64 sums; J in fp64",
                           "bf16x6": "three-piece bf16 split of f32, six products, fp64 sums; J in fp64",
                           "bf16x3": "two-piece bf16 split, three products, fp64 sums; J in fp64",
                           "fp64": "f64"}[prec],
            "dE_vs_fp64_Eh": acc[prec]["dE"],
            "max_abs_dJ": acc[prec]["max_abs_dJ"],
            "max_abs_dK": acc[prec]["max_abs_dK"],
            "fp64_builds_per_s": round(rate64, 2),
            "speedup_vs_fp64": round(rate / rate64, 3),
            "host_path_builds_per_s": round(host_rate, 2),
            "int8_shares_read": {k: round(v, 4) for k, v in zero_blocks.items()},
            "modes": {p: {"builds_per_s": round(runs[p][0], 2), "dE_vs_fp64_Eh": acc[p]["dE"],
                          "max_abs_dK": acc[p]["max_abs_dK"],
                          "phase_ms": {k: round(v, 4) for k, v in runs[p][2].items()},
                          "split_counts": tuning[p],
                          "kernels": {k: (None if v is None else {kk: (round(vv, 4) if isinstance(vv, float) else vv)
                                                                  for kk, vv in v.items()})
                                      for k, v in kern[p].items()}} for p in modes},
            "perturbed_densities": pert,
            "ragged_batch": ragged,
            "E_rhf_fp64": res.E,
            "setup_s": {"host_integrals": round(t_ints, 1), "B_device": round(t_B, 2), "scf": round(t_scf, 1)},
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
