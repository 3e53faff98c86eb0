#!/usr/bin/env python3
"""Fock-build benchmark on one MI355X.

Workload (headline): 64 waters (seeded orientations in a 12.4 A box),
cc-pVDZ (M = 1600 Cartesian) + cc-pVDZ-RIFIT (Maux = 6144), nocc = 320, RHF
with RI-JK in precision=mixed.  A step is one Fock build F = H + J - K/2 on
a new set of occupied orbitals (an SCF-like sequence whose occupied space
moves by 1e-3 x 0.7^k), with B already resident in HBM: after the first
build each one forms K and J as increments on the fp16 matrix cores, and
a full FP64 build follows after at most 16 increments or once the summed
orbital change reaches the default budget, so the timed loop carries the
refreshes at their natural rate.

Prints one JSON line: builds/s, the roofline of the dominant kernel (timed
live with HIP events on the build stream), the FP64 builds of the same
workload (occupied-form RI, the reference's density-form RI, the
four-centre direct build), mixed vs FP64 RI SCF (energy difference and
wall time, the RI gradient), the C10H22/cc-pVDZ stored four-centre build,
the C24H12 and C10H22/6-31G* gradient legs, and a CPU baseline: the FP64
host oracle's direct J/K build of the same 64 waters on a bounded sample,
extrapolated.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (scaled x2.0 / x1.0 by a 2 GiB copy), keyed by (workload, kernel); the
# commands and raw counts are in profiles/stored_digest_c10h22_ccpvdz.md and
# profiles/ri_mixed_64water_ccpvdz.md.
MEASURED_TRAFFIC = {
    ("C10H22/cc-pVDZ", "fock_slab(pp|ps) stored-digest"): 274.7e6,
    # profiles/ri_mixed_64water_ccpvdz.md (FETCH_SIZE x2.0, WRITE_SIZE x1.0 by a 2 GiB copy)
    ("64 waters", "k_rik_dk (Z = T_A T_E^T, fp16 MFMA)"): 143.0e9,
    ("64 waters", "k_rik_tstep (T_E = B E, fp16 MFMA)"): 118.3e9,
}
TRAFFIC_SOURCE = "stored constant: rocprofv3 FETCH_SIZE / WRITE_SIZE per launch, scaled by a 2 GiB copy (profiles/)"
FP16_MFMA_PEAK_TFS = 2500.0  # MI355X dense fp16 MFMA, TFLOP/s
# kernel ids of the engine's profile rows (incremental RI build)
RIK_NAMES = {80020: "align (Procrustes)", 80021: "split E", 80022: "k_rik_tstep (T_E = B E, fp16 MFMA)",
             80023: "dW", 80024: "J increment (B planes)", 80025: "k_rik_dk (Z = T_A T_E^T, fp16 MFMA)",
             80026: "K reduce (FP64)", 80005: "T=B*X dgemm", 80007: "W from T", 80008: "K=T*T^T dgemm",
             80002: "J (FP64 B)"}


FP64_VECTOR_PEAK_TFS = 78.6  # MI355X FP64 vector, TFLOP/s


def md_algorithmic_rate(eng, build_s, table="classes"):
    """Algorithmic FP64 work of a direct build: per quartet class the primitive
    quartets that pass the Schwarz screen (the engine's count) times the FP64
    operations of one primitive quartet in the oracle's McMurchie-Davidson
    routine (tests/golden/md_flops.json, counted by running that routine on
    op-counting numbers), over the build time.  The kernels factor the
    Hermite contraction (E tables and bra intermediates shared across
    components), so they execute fewer operations than this count; the rate
    is the reference algorithm's work per second.  table="gradient_classes":
    the nuclear-derivative routine's count, for the two-electron gradient."""
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "md_flops.json")))[table]
    cs = eng.class_stats()
    flops = sum(c["primitive"] * fx["%d %d %d %d" % c["L"]][0] for c in cs)
    prim = sum(c["primitive"] for c in cs)
    tfs = flops / build_s / 1e12
    return dict(primitive_quartets=prim, contracted_quartets=sum(c["contracted"] for c in cs),
                reference_md_flops=flops, reference_md_tflops=round(tfs, 2),
                frac_of_fp64_vector_peak=round(tfs / FP64_VECTOR_PEAK_TFS, 3),
                primitive_quartets_per_s=round(prim / build_s, 1), unit_note="reference-MD operation count")


def orbital_sequence(M, n, seed, amp=1e-3, decay=0.7):
    """n orthonormal M x M orbital matrices (torch, on the GPU): a random
    orthogonal C_0 (host QR), then C_k = C_(k-1) exp(A_k) with A_k a random
    antisymmetric matrix of norm ~ amp decay^k (exponential by its Taylor
    series, matrix products on the GPU), so that the occupied space moves
    by about amp decay^k per step, as between SCF diagonalisations."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    C = torch.tensor(np.linalg.qr(rng.normal(size=(M, M)))[0], dtype=torch.float64, device="cuda").contiguous()
    g = torch.Generator(device="cuda").manual_seed(seed)
    out = [C]
    for k in range(1, n):
        A = torch.randn(M, M, generator=g, device="cuda", dtype=torch.float64) * (amp * decay ** (k - 1) / M)
        A = A - A.T
        R = torch.eye(M, dtype=torch.float64, device="cuda")
        term = R.clone()
        for j in range(1, 8):
            term = term @ A / j
            R = R + term
        C = (C @ R).contiguous()
        out.append(C)
    return out


def ri_headline(steps, warmup, log_fn):
    """The headline: precision=mixed RI-JK builds of the 64-water workload on
    a moving orbital sequence (incremental fp16-MFMA builds plus the FP64
    refreshes the default policy makes)."""
    import numpy as np
    import torch
    from dft_gansu_amd import FockEngine
    from dft_gansu_amd import molecule as gm
    mol, aux = gm.config_molecule(4)
    M, nocc = mol.n_basis, mol.n_electrons // 2
    nprof = 3
    seq = orbital_sequence(M, warmup + steps + nprof + 1, seed=5)
    dens = [(2.0 * C[:, :nocc] @ C[:, :nocc].T).contiguous() for C in seq]
    H = torch.zeros((M, M), dtype=torch.float64, device="cuda")
    F = torch.empty_like(H)
    stream = torch.cuda.current_stream()
    e = FockEngine(mol, aux=aux, eri_method="ri", precision="mixed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.precomputation()
    torch.cuda.synchronize()
    pre = time.perf_counter() - t0
    log_fn(f"[bench] 64 waters mixed RI: precompute {pre:.2f} s")
    k = 0
    for _ in range(warmup):
        e.compute_fock_matrix_occ(dens[k], seq[k], nocc, H, F, stream=stream)
        k += 1
    n_inc0 = e.stats()["ri_incremental_builds"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        e.compute_fock_matrix_occ(dens[k], seq[k], nocc, H, F, stream=stream)
        k += 1
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n_inc = e.stats()["ri_incremental_builds"] - n_inc0
    log_fn(f"[bench] {steps} mixed builds in {elapsed:.3f} s ({n_inc} incremental, {steps - n_inc} FP64)")
    # per-kernel rows of incremental builds (HIP events on the build stream)
    agg = {}
    nrows = 0
    for _ in range(nprof):
        before = e.stats()["ri_incremental_builds"]
        e.set_profiling(True)
        e.compute_fock_matrix_occ(dens[k], seq[k], nocc, H, F, stream=stream)
        k += 1
        rows = e.last_profile()
        e.set_profiling(False)
        if e.stats()["ri_incremental_builds"] == before:
            continue  # a refresh build: not an incremental profile
        nrows += 1
        for kid, ms, by, fl, un in rows:
            a = agg.setdefault(int(kid), [0.0, by, fl, un, 0])
            a[0] += ms
            a[4] += 1
    for v in agg.values():
        v[0] /= max(1, v[4])  # per launch
    kernels = sorted(((RIK_NAMES.get(kid, str(kid)), v[0], v[1], v[2]) for kid, v in agg.items()), key=lambda r: -r[1])
    dom = kernels[0]
    tfs = dom[3] / (dom[1] * 1e-3) / 1e12
    roof = dict(bound="mfma", kernel=dom[0], achieved=round(tfs, 1), peak=FP16_MFMA_PEAK_TFS, unit="TFLOP/s",
                frac=round(tfs / FP16_MFMA_PEAK_TFS, 4), traffic=MEASURED_TRAFFIC.get(("64 waters", dom[0])),
                traffic_source=TRAFFIC_SOURCE,
                flops_per_launch=dom[3], avg_launch_ms=round(dom[1], 3),
                flops_note="three fp16 MFMA products (hi/lo split) per multiply-add of the algorithm")
    per_kernel = []
    for name, ms, by, fl in kernels:
        row = dict(kernel=name, ms=round(ms, 3))
        if fl > 0:
            row.update(tflops=round(fl / (ms * 1e-3) / 1e12, 1), frac_fp16_peak=round(fl / (ms * 1e-3) / 1e12 /
                                                                                     FP16_MFMA_PEAK_TFS, 3))
        if by > 0:
            row.update(GBs=round(by / (ms * 1e-3) / 1e9, 1))
        per_kernel.append(row)
    log_fn(f"[bench] dominant kernel {dom[0]}: {dom[1]:.3f} ms, {tfs:.0f} TFLOP/s")
    del e, seq, dens
    torch.cuda.empty_cache()
    cfgd = dict(workload="64 H2O cc-pVDZ + cc-pVDZ-RIFIT RHF RI-JK, precision=mixed (fp16-MFMA incremental K/J, "
                         "FP64 refresh by the default policy)", n_basis_cartesian=M, n_aux_cartesian=aux.n_basis,
                nocc=nocc, eri_method="ri", precision="mixed", precompute_s=round(pre, 3),
                incremental_builds_timed=n_inc, fp64_builds_timed=steps - n_inc)
    return dict(value=round(steps / elapsed, 3), ms_per_step=round(elapsed / steps * 1e3, 3), roofline=roof,
                kernels=per_kernel, config=cfgd)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_name(kid):
    kid = int(kid)
    f32 = kid >= 1000000  # class with an FP32-stored part (precision=mixed)
    kid %= 1000000
    kind = {1: "fock_slab", 3: "jket"}.get(kid // 100000, "eri_class") + ("[fp64+fp32 store]" if f32 else "")
    kid %= 100000
    mode = {0: "direct", 1: "store", 2: "stored-digest"}.get(kid // 10000, "?")
    l = "spdf"
    c = kid % 10000
    la, lb, lc, ld = c // 1000, (c // 100) % 10, (c // 10) % 10, c % 10
    return f"{kind}({l[la]}{l[lb]}|{l[lc]}{l[ld]}) {mode}"


def cpu_baseline(cfg, target_s):
    """FP64 host oracle direct J/K on a strided sample of bra primitive-pair rows."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from dft_gansu_amd import molecule as gm
    c = gm.CONFIGS[cfg]
    Z, xyz = c["geom"]()
    p, t, n = O.build_molecule(gm.basis_path(c["basis"]), Z, xyz)
    Q = O.schwarz(p, t)
    rng = np.random.default_rng(5)
    C = rng.normal(size=(len(n), max(1, len(n) // 3))) / np.sqrt(len(n))
    D = 2.0 * C @ C.T
    stride = max(64, O.n_pairs(t) // 32)
    while True:
        t0 = time.perf_counter()
        O.jk(p, t, n, Q, D, 1e-12, sample_stride=stride)
        dt = time.perf_counter() - t0
        if dt * 4 > target_s or stride == 1:
            break
        stride = max(1, int(stride * max(dt, 1e-3) / target_s * 0.9)) if dt > 0.05 else max(1, stride // 8)
    full_s = dt * stride
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
    except OSError:
        pass
    return dict(value=1.0 / full_s, unit="Fock builds/s", cores=O.lib().orc_num_threads(), kind="port",
                nproc=os.cpu_count(), cpu_model=model, omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                sample=f"{c['name']}: the FP64 host oracle's four-centre direct J/K (MD integrals + digestion, "
                       f"Schwarz 1e-12) over every {stride}-th bra primitive-pair row ({dt:.1f} s measured, "
                       f"x{stride} extrapolated; full build ~{full_s:.0f} s); the host cannot hold the RI tensor "
                       f"(126 GB FP64), so the GPU's RI build is set against the CPU's four-centre build")


def ri_leg(log_fn, mixed_rate, no_scf=False):
    """64-water cc-pVDZ + cc-pVDZ-RIFIT: the FP64 builds the mixed headline is
    measured against (four-centre direct build, the drop-in; the reference's
    density-form RI-K; the occupied-form FP64 RI), the RI gradient, and the
    mixed RI SCF (with its analytic gradient) against the FP64 RI SCF."""
    import numpy as np
    import torch
    from dft_gansu_amd import FockEngine, Parameters, rhf, rhf_gradient
    from dft_gansu_amd import molecule as gm
    mol, aux = gm.config_molecule(4)
    M, nocc = mol.n_basis, mol.n_electrons // 2
    rng = np.random.default_rng(5)
    C = torch.tensor(np.linalg.qr(rng.normal(size=(M, M)))[0], dtype=torch.float64, device="cuda")
    D = (2.0 * C[:, :nocc] @ C[:, :nocc].T).contiguous()
    H = torch.zeros_like(D)
    out = {"workload": "64 H2O cc-pVDZ (M=%d) + cc-pVDZ-RIFIT (Maux=%d), nocc=%d" % (M, aux.n_basis, nocc)}

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n

    e = FockEngine(mol, eri_method="direct")
    e.precomputation()
    out["fp64_direct_build_s"] = round(timed(lambda: e.compute_fock_matrix(D, H), 2), 4)
    del e
    log_fn(f"[bench] 64-water FP64 direct build {out['fp64_direct_build_s']} s")
    for name, kw in (("ri_fp64_density", dict(precision="fp64", ri_k="density")),
                     ("ri_fp64_occupied", dict(precision="fp64", ri_k="occupied"))):
        e = FockEngine(mol, aux=aux, eri_method="ri", **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.precomputation()
        torch.cuda.synchronize()
        pre = time.perf_counter() - t0
        if name == "ri_fp64_density":
            t = timed(lambda: e.compute_fock_matrix(D, H), 1)
        else:
            t = timed(lambda: e.compute_fock_matrix_occ(D, C, nocc, H), 3)
        out[name + "_build_s"] = round(t, 4)
        out[name + "_precompute_s"] = round(pre, 3)
        if name != "ri_fp64_density":  # per-phase rates of one build from the engine's HIP-event rows
            e.set_profiling(True)
            e.compute_fock_matrix_occ(D, C, nocc, H)
            rows = e.last_profile()
            e.set_profiling(False)
            phases = {}
            for kid, label, unit in ((80005, "T=B*X dgemm", "TFLOP/s"), (80008, "K=T*T^T dgemm", "TFLOP/s"),
                                     (80002, "J (FP64 B)", "GB/s")):
                sel = rows[rows[:, 0] == kid]
                if len(sel) == 0:
                    continue
                ms = float(sel[:, 1].sum())
                if unit == "TFLOP/s":
                    a = float(sel[:, 3].sum()) / (ms * 1e-3) / 1e12
                    phases[label] = dict(ms=round(ms, 3), achieved=round(a, 2), peak=78.6, unit=unit,
                                         frac=round(a / 78.6, 3))
                else:
                    a = float(sel[:, 2].sum()) / (ms * 1e-3) / 1e9
                    phases[label] = dict(ms=round(ms, 3), achieved=round(a, 1), peak=HBM_PEAK_GBS, unit=unit,
                                         frac=round(a / HBM_PEAK_GBS, 3))
            out[name + "_phases"] = phases
            log_fn(f"[bench] 64-water {name} phases: {phases}")
        if name == "ri_fp64_occupied":  # RI-JK energy gradient for the same density and orbitals
            g = [None]

            def grad():
                g[0] = e.gradient_2e(D, C=C, nocc=nocc)

            out["ri_gradient_2e_s"] = round(timed(grad, 1), 4)
            out["ri_gradient_translational_residual"] = float(g[0].sum(dim=0).abs().max())
            log_fn(f"[bench] 64-water RI two-electron gradient {out['ri_gradient_2e_s']} s")
        del e
        torch.cuda.empty_cache()
        log_fn(f"[bench] 64-water {name} build {t:.4f} s (precompute {pre:.2f} s)")
    mixed_s = 1.0 / mixed_rate
    out["ri_mixed_build_s"] = round(mixed_s, 4)
    out["speedup_mixed_ri_vs_fp64_direct"] = round(out["fp64_direct_build_s"] / mixed_s, 2)
    out["speedup_mixed_ri_vs_fp64_ri_density"] = round(out["ri_fp64_density_build_s"] / mixed_s, 2)
    out["speedup_mixed_ri_vs_fp64_ri_occupied"] = round(out["ri_fp64_occupied_build_s"] / mixed_s, 2)
    if no_scf:
        return out
    scf = {}
    # each SCF in a fresh process, as a user's run would be: in a process
    # that has already freed a buffer of B's size (126 GB), allocating one
    # again costs ~2 s while the driver clears it
    # (and a process started right after one that freed B and its planes
    # waits for the driver to clear them: measured 5.0 s instead of 1.5 s of
    # precompute; after a 30 s pause the next one precomputed in 1.4 s, so
    # the legs are spaced to time each on a device the last one has left)
    for i, (name, kw, grad) in enumerate((("fp64", dict(ri_k="occupied"), False),
                                          ("mixed", dict(precision="mixed"), False),
                                          ("mixed_with_gradient", dict(precision="mixed"), True))):
        if i > 0:
            log_fn("[bench] 15 s pause: the driver clears the memory the previous SCF process freed")
            time.sleep(15.0)
        r = scf_in_fresh_process(4, dict(eri_method="ri", convergence_energy_threshold=1e-9, maxiter=60, **kw), grad)
        scf[name] = r
        log_fn(f"[bench] 64-water RI SCF {name} (fresh process): {scf[name]}")
    scf["note"] = ("each leg a fresh process, 15 s apart: a process started right after one that freed B "
                   "(126 GB) and its planes waits for the driver to clear that memory at its first allocation "
                   "(precompute 5.0 s instead of 1.5 s, measured); the pause lets the clearing finish first")
    scf["dE_mixed_vs_fp64"] = scf["mixed"]["energy"] - scf["fp64"]["energy"]
    scf["speedup_mixed_vs_fp64"] = round(scf["fp64"]["wall_s"] / scf["mixed"]["wall_s"], 2)
    out["scf"] = scf
    return out


def coronene_leg(log_fn):
    """C24H12/cc-pVDZ + cc-pVDZ-RIFIT: every four-centre class up to (dd|dd)
    through the batched class launches (stored FP64, stored mixed, direct) and
    the RI-JK build (FP64 occupied form vs mixed), then the stored and RI SCF in
    FP64 and mixed with their energy differences."""
    import numpy as np
    import torch
    from dft_gansu_amd import FockEngine, Parameters, rhf
    from dft_gansu_amd import molecule as gm
    mol, aux = gm.config_molecule(2)
    M, nocc = mol.n_basis, mol.n_electrons // 2
    rng = np.random.default_rng(7)
    C = torch.tensor(np.linalg.qr(rng.normal(size=(M, M)))[0], dtype=torch.float64, device="cuda")
    D = (2.0 * C[:, :nocc] @ C[:, :nocc].T).contiguous()
    H = torch.zeros_like(D)
    out = {"workload": "C24H12 cc-pVDZ (M=%d) + cc-pVDZ-RIFIT (Maux=%d), nocc=%d" % (M, aux.n_basis, nocc)}

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n

    legs = (("stored_fp64", dict(eri_method="stored"), False), ("stored_mixed", dict(eri_method="stored",
            precision="mixed"), False), ("direct_fp64", dict(eri_method="direct"), False),
            ("ri_fp64", dict(eri_method="ri", ri_k="occupied"), True), ("ri_mixed", dict(eri_method="ri",
            precision="mixed"), True))
    for name, kw, ri in legs:
        e = FockEngine(mol, aux=aux if ri else None, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.precomputation()
        torch.cuda.synchronize()
        pre = time.perf_counter() - t0
        t = timed((lambda: e.compute_fock_matrix_occ(D, C, nocc, H)) if ri else (lambda: e.compute_fock_matrix(D, H)),
                  5)
        out[name] = dict(build_s=round(t, 5), builds_per_s=round(1.0 / t, 2), precompute_s=round(pre, 3))
        if not ri:
            st = e.stats()
            out[name]["stored_bytes"] = st["stored_bytes"]
            out[name]["contracted_quartets"] = st["contracted_quartets"]
            if name == "direct_fp64":
                out[name]["md"] = md_algorithmic_rate(e, t)
        del e
        torch.cuda.empty_cache()
        log_fn(f"[bench] C24H12 {name}: {out[name]}")
    out["speedup_ri_mixed_vs_direct_fp64"] = round(out["direct_fp64"]["build_s"] / out["ri_mixed"]["build_s"], 2)
    scf = {}
    for name, kw, ri in (("stored_fp64", dict(eri_method="stored"), False),
                         ("stored_fp64_purify", dict(eri_method="stored", scf_solver="purify"), False),
                         ("stored_mixed", dict(eri_method="stored", precision="mixed"), False),
                         ("ri_fp64", dict(eri_method="ri", ri_k="occupied"), True),
                         ("ri_mixed", dict(eri_method="ri", precision="mixed"), True)):
        t0 = time.perf_counter()
        r = rhf(mol, Parameters(convergence_energy_threshold=1e-8, maxiter=100, **kw), aux=aux if ri else None)
        scf[name] = dict(wall_s=round(time.perf_counter() - t0, 3), energy=r["e_total"], iterations=r["iterations"],
                         converged=r["converged"])
        log_fn(f"[bench] C24H12 SCF {name}: {scf[name]}")
    scf["dE_stored_mixed_vs_fp64"] = scf["stored_mixed"]["energy"] - scf["stored_fp64"]["energy"]
    scf["dE_stored_purify_vs_diagonalize"] = scf["stored_fp64_purify"]["energy"] - scf["stored_fp64"]["energy"]
    scf["dE_ri_mixed_vs_fp64"] = scf["ri_mixed"]["energy"] - scf["ri_fp64"]["energy"]
    out["scf"] = scf
    return out


def gradient_leg(log_fn):
    """C10H22/6-31G*: SCF then the analytic gradient; gradient s/call is the
    backward work alone (two-electron + one-electron kernels) at the converged
    density."""
    import numpy as np
    import torch
    from dft_gansu_amd import FockEngine, Parameters, oneint_gradient, rhf, rhf_gradient
    from dft_gansu_amd import molecule as gm
    mol, _ = gm.config_molecule(3)
    prm = Parameters(convergence_energy_threshold=1e-11, maxiter=200)
    t0 = time.perf_counter()
    r = rhf_gradient(mol, prm)
    wall = time.perf_counter() - t0
    D = torch.tensor(rhf(mol, prm)["D"], dtype=torch.float64, device="cuda")
    W = (0.5 * D @ D).contiguous()  # any symmetric W: the kernel cost does not depend on it
    eng = FockEngine(mol, eri_method="direct")
    eng.precomputation()
    eng.gradient_2e(D)
    oneint_gradient(mol, D, W)
    torch.cuda.synchronize()
    n = 3
    t0 = time.perf_counter()
    for _ in range(n):
        eng.gradient_2e(D)
        oneint_gradient(mol, D, W)
    torch.cuda.synchronize()
    per = (time.perf_counter() - t0) / n
    # two-electron part alone, for the algorithmic derivative-MD rate
    t0 = time.perf_counter()
    for _ in range(n):
        eng.gradient_2e(D)
    torch.cuda.synchronize()
    per2 = (time.perf_counter() - t0) / n
    md = md_algorithmic_rate(eng, per2, table="gradient_classes")
    del eng
    g = r["grad"]
    out = dict(workload=f"C10H22 6-31G* RHF (M={mol.n_basis}, {mol.n_atoms} atoms)", gradient_s_per_call=round(per, 4),
               gradient_2e_s=round(per2, 4), gradient_2e_md=md,
               scf_plus_gradient_wall_s=round(wall, 3), energy=r["e_total"],
               max_abs_gradient=float(np.max(np.abs(g))),
               translational_residual=float(np.max(np.abs(g.sum(axis=0)))))
    log_fn(f"[bench] gradient C10H22/6-31G*: {per * 1e3:.1f} ms/call, SCF+gradient {wall:.2f} s")
    return out


def direct_leg(cfgs, log_fn, steps=3):
    """Direct (integral-recomputing) FP64 four-centre builds of the given
    workloads: wall time per build, then one profiled build whose per-class
    kernel rows (HIP events, the classes run in sequence) are reported."""
    import numpy as np
    import torch
    from dft_gansu_amd import FockEngine
    from dft_gansu_amd import molecule as gm
    out = {}
    for cfg in cfgs:
        mol, _ = gm.config_molecule(cfg)
        M, nocc = mol.n_basis, mol.n_electrons // 2
        rng = np.random.default_rng(7)
        C = torch.tensor(np.linalg.qr(rng.normal(size=(M, M)))[0], dtype=torch.float64, device="cuda")
        D = (2.0 * C[:, :nocc] @ C[:, :nocc].T).contiguous()
        H = torch.zeros_like(D)
        e = FockEngine(mol, eri_method="direct")
        e.precomputation()
        e.compute_fock_matrix(D, H)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            e.compute_fock_matrix(D, H)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        e.set_profiling(True)
        e.compute_fock_matrix(D, H)
        torch.cuda.synchronize()
        rows = e.last_profile()
        e.set_profiling(False)
        top = sorted(rows, key=lambda r: -r[1])[:12]
        name = {1: "C10H22/cc-pVDZ", 2: "C24H12/cc-pVDZ"}.get(cfg, f"config {cfg}")
        out[name] = dict(M=M, ms_per_build=round(ms, 2), sequential_kernel_ms=round(float(rows[:, 1].sum()), 2),
                         classes={kernel_name(r[0]): dict(ms=round(float(r[1]), 3), quartets=int(r[4])) for r in top})
        log_fn(f"[bench] direct {name} (M={M}): {ms:.2f} ms/build")
        del e
        torch.cuda.empty_cache()
    return out


def stored_leg(cfg, method, precision, steps, warmup, no_direct, no_scf, no_mixed, log_fn):
    """Four-centre leg (C10H22/cc-pVDZ by default): the stored (or direct)
    build with the integrals resident in HBM, the live roofline of its
    dominant digestion kernel, the direct build rate, the SCF, and the mixed
    store (FP32 for quartets bounded below 1e-6)."""
    import numpy as np
    import torch
    from dft_gansu_amd import FockEngine, Parameters, oneint, rhf
    from dft_gansu_amd import molecule as gm
    dev = torch.device("cuda:0")
    cfgd = gm.CONFIGS[cfg]
    mol, aux = gm.config_molecule(cfg)
    M = mol.n_basis
    eng = FockEngine(mol, eri_method=method, precision=precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.precomputation()
    torch.cuda.synchronize()
    precompute_s = time.perf_counter() - t0
    st = eng.stats()
    S, T, V = oneint(mol)
    H = (T + V).contiguous()
    rng = np.random.default_rng(5)
    Cm = rng.normal(size=(M, mol.n_electrons // 2)) / np.sqrt(M)
    D = torch.tensor(2.0 * Cm @ Cm.T, dtype=torch.float64, device=dev)
    F = torch.empty_like(D)
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        eng.compute_fock_matrix(D, H, F, stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.compute_fock_matrix(D, H, F, stream=stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # live per-kernel timing (HIP events on the build stream), averaged over builds
    eng.set_profiling(True)
    agg = {}
    nprof = max(3, min(10, steps))
    for _ in range(nprof):
        eng.compute_fock_matrix(D, H, F, stream=stream)
        for kid, ms, by, fl, un in eng.last_profile():
            a = agg.setdefault(int(kid), [0.0, by, fl, un])
            a[0] += ms / nprof
    eng.set_profiling(False)
    torch.cuda.synchronize()
    kernels = sorted(((k, v[0], v[1], v[2], v[3]) for k, v in agg.items()), key=lambda r: -r[1])
    dom = kernels[0]
    build_kernel_ms = sum(r[1] for r in kernels)
    # algorithmic bytes of a build: every stored integral once (the ket-major J
    # pass re-reads them; that is extra traffic, not algorithmic work)
    all_bytes = sum(r[2] for r in kernels if (int(r[0]) % 1000000) // 100000 != 3)
    if method == "stored":
        gbs = dom[2] / (dom[1] * 1e-3) / 1e9
        roof = dict(bound="hbm", kernel=kernel_name(dom[0]), achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(gbs / HBM_PEAK_GBS, 4), traffic=MEASURED_TRAFFIC.get((cfgd["name"], kernel_name(dom[0]))),
                    traffic_source=TRAFFIC_SOURCE,
                    bytes_per_launch=dom[2], avg_launch_ms=round(dom[1], 4),
                    all_digest_kernels=dict(ms=round(build_kernel_ms, 4), bytes=all_bytes,
                                            GBs=round(all_bytes / (build_kernel_ms * 1e-3) / 1e9, 1)))
    else:
        roof = dict(bound="fp64-valu", kernel=kernel_name(dom[0]), avg_launch_ms=round(dom[1], 4))
    out = dict(workload=f"{cfgd['name']} RHF {method} ERI, Schwarz 1e-12", n_basis_cartesian=M,
               fock_builds_per_s=round(steps / elapsed, 3), ms_per_build=round(elapsed / steps * 1e3, 4),
               precompute_s=round(precompute_s, 4), contracted_quartets=st["contracted_quartets"],
               stored_bytes=st["stored_bytes"], roofline=roof,
               kernels=[dict(kernel=kernel_name(k), ms=round(ms, 4), quartets=un) for k, ms, by, fl, un in kernels[:8]])
    log_fn(f"[bench] {cfgd['name']} {method}: {out['ms_per_build']} ms/build, dominant {roof['kernel']}")
    del eng
    torch.cuda.empty_cache()
    if not no_direct and method != "direct":
        e2 = FockEngine(mol, eri_method="direct", precision=precision)
        e2.precomputation()
        e2.compute_fock_matrix(D, H, F, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nd = 3
        for _ in range(nd):
            e2.compute_fock_matrix(D, H, F, stream=stream)
        torch.cuda.synchronize()
        out["direct_fock_builds_per_s"] = round(nd / (time.perf_counter() - t0), 3)
        log_fn(f"[bench] direct build rate {out['direct_fock_builds_per_s']} /s")
        out["direct_md"] = md_algorithmic_rate(e2, 1.0 / out["direct_fock_builds_per_s"])
        log_fn(f"[bench] direct build, algorithmic MD work: {out['direct_md']}")
        del e2
        if not no_scf:
            # direct SCF with full builds and with delta-density builds (a full one every 8)
            dscf = {}
            for name, kw in (("full_builds", {}), ("incremental_fock_8", dict(incremental_fock=8))):
                t0 = time.perf_counter()
                rd = rhf(mol, Parameters(eri_method="direct", precision=precision, convergence_energy_threshold=1e-8,
                                         maxiter=100, **kw))
                dscf[name] = dict(wall_s=round(time.perf_counter() - t0, 3), energy=rd["e_total"],
                                  iterations=rd["iterations"], converged=rd["converged"])
            dscf["dE_incremental_vs_full"] = dscf["incremental_fock_8"]["energy"] - dscf["full_builds"]["energy"]
            out["direct_scf"] = dscf
            log_fn(f"[bench] direct SCF: {dscf}")
    if not no_scf:
        t0 = time.perf_counter()
        r = rhf(mol, Parameters(eri_method=method, precision=precision, convergence_energy_threshold=1e-8,
                                maxiter=100), aux=aux)
        out.update(scf_wall_s=round(time.perf_counter() - t0, 3), scf_energy=r["e_total"],
                   scf_iterations=r["iterations"], scf_converged=r["converged"])
        log_fn(f"[bench] SCF {out['scf_wall_s']} s, E = {r['e_total']:.10f} ({r['iterations']} it)")
        if method != "ri":
            # the device-resident iteration: one captured graph per iteration
            t0 = time.perf_counter()
            rp = rhf(mol, Parameters(eri_method=method, precision=precision, convergence_energy_threshold=1e-8,
                                     maxiter=100, scf_solver="purify"), aux=aux)
            out["scf_purify"] = dict(wall_s=round(time.perf_counter() - t0, 3), energy=rp["e_total"],
                                     iterations=rp["iterations"], graph_launches=rp["graph_launches"],
                                     converged=rp["converged"], dE_vs_diagonalize=rp["e_total"] - r["e_total"])
            log_fn(f"[bench] SCF purify (graph per iteration): {out['scf_purify']}")
    if not no_mixed and precision == "fp64" and method == "stored":
        em = FockEngine(mol, eri_method="stored", precision="mixed")
        em.precomputation()
        for _ in range(warmup):
            em.compute_fock_matrix(D, H, F, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            em.compute_fock_matrix(D, H, F, stream=stream)
        torch.cuda.synchronize()
        tm = (time.perf_counter() - t0) / steps
        sm = em.stats()
        del em
        torch.cuda.empty_cache()
        mixed = dict(fock_builds_per_s=round(1.0 / tm, 3), ms_per_build=round(tm * 1e3, 4),
                     stored_bytes=sm["stored_bytes"], fp32_quartets=sm["fp32_quartets"], fp32_bound_threshold=1e-6)
        if not no_scf:
            t0 = time.perf_counter()
            r = rhf(mol, Parameters(eri_method="stored", precision="mixed", convergence_energy_threshold=1e-8,
                                    maxiter=100))
            mixed.update(scf_wall_s=round(time.perf_counter() - t0, 3), scf_energy=r["e_total"],
                         scf_iterations=r["iterations"])
            if "scf_energy" in out:
                mixed["dE_vs_fp64"] = r["e_total"] - out["scf_energy"]
        out["mixed"] = mixed
        log_fn(f"[bench] mixed store: {mixed}")
    return out


def scf_in_fresh_process(cfg, params, gradient=False, timeout=900):
    """One RHF SCF (optionally with its analytic gradient) of a benchmark
    workload in a child process; returns the child's result dict (wall time
    of the SCF call, energy, iterations)."""
    import subprocess
    job = json.dumps(dict(cfg=cfg, params=params, gradient=gradient))
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--scf-child", job], capture_output=True,
                       text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"SCF child failed ({r.returncode}): {r.stderr[-2000:]}")
    return json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])


def scf_child(job):
    import numpy as np
    from dft_gansu_amd import Parameters, rhf, rhf_gradient
    from dft_gansu_amd import molecule as gm
    job = json.loads(job)
    mol, aux = gm.config_molecule(job["cfg"])
    prm = Parameters(**job["params"])
    use_aux = aux if job["params"].get("eri_method") == "ri" else None
    t0 = time.perf_counter()
    r = (rhf_gradient if job["gradient"] else rhf)(mol, prm, aux=use_aux)
    out = dict(wall_s=round(time.perf_counter() - t0, 3), energy=r["e_total"], iterations=r["iterations"],
               converged=r["converged"])
    if job["gradient"]:
        out["max_abs_grad"] = float(np.max(np.abs(r["grad"])))
    print(json.dumps(out), flush=True)


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "--scf-child":
        scf_child(sys.argv[2])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=18)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-scf", action="store_true", help="skip the SCF runs of every leg")
    ap.add_argument("--no-ri", action="store_true", help="skip the 64-water FP64 comparison leg")
    ap.add_argument("--no-stored", action="store_true", help="skip the C10H22/cc-pVDZ four-centre leg")
    ap.add_argument("--no-grad", action="store_true", help="skip the C10H22/6-31G* gradient leg")
    ap.add_argument("--no-coronene", action="store_true", help="skip the C24H12 four-centre + RI-JK leg")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="run one device copy of this many GiB first (a known-size kernel that scales the "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counts of the other kernels)")
    ap.add_argument("--direct-only", action="store_true",
                    help="only the direct FP64 four-centre builds of C10H22 and C24H12 (cc-pVDZ), "
                         "with per-class kernel times")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the Fock build runs on a single MI355X in this build "
              "(multi-GPU sharding is described, not run); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from dft_gansu_amd import Parameters, rhf
    from dft_gansu_amd import molecule as gm

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    # one small SCF first (water STO-3G), so that everything below runs with
    # rocBLAS / rocSOLVER already initialised in this process
    rhf(gm.config_molecule(0)[0], Parameters(convergence_energy_threshold=1e-8))
    if args.direct_only:
        print(json.dumps({"direct_fp64": direct_leg([1, 2], log)}), flush=True)
        return
    if args.copy_probe_gib > 0:
        n = int(args.copy_probe_gib * 2 ** 30 / 8)
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        torch.cuda.empty_cache()
        log(f"[bench] copy probe: {args.copy_probe_gib} GiB read, {args.copy_probe_gib} GiB written")
    head = ri_headline(args.steps, args.warmup, log)
    extra = {}
    if not args.no_ri:
        extra["ri_64water"] = ri_leg(log, head["value"], no_scf=args.no_scf)
        torch.cuda.empty_cache()
    if not args.no_stored:
        extra["c10h22_stored"] = stored_leg(1, "stored", "fp64", 20, 3, False, args.no_scf, False, log)
        torch.cuda.empty_cache()
    if not args.no_coronene:
        extra["c24h12"] = coronene_leg(log)
        torch.cuda.empty_cache()
    if not args.no_grad:
        extra["gradient"] = gradient_leg(log)
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle) ...")
        cpu = cpu_baseline(4, args.cpu_seconds)
        log(f"[bench] CPU baseline {cpu['value']:.3e} builds/s on {cpu['cores']} threads")

    out = {
        "metric": "Fock builds/s + SCF wall s at |dE|<=1uEh, mixed vs FP64 on GPU; gradient s/call",
        "value": head["value"],
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f16",
        "data": "synthetic (seeded 64-water geometry; SCF-like sequence of random orthonormal orbitals)",
        "config": head["config"],
        "roofline": head["roofline"],
        "kernels": head["kernels"],
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
