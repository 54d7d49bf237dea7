"""Fock-build benchmark: (H2O)64 / cc-pVDZ (1536 basis functions) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--system h2o64|h2o16|c20h42]
                    [--no-scf] [--no-cpu] [--no-mixed] [--no-grad] [--no-atomics]
                    [--no-c5] [--density P.npy] [--save-density P.npy]

One step is one full J/K build (screened task list + ERI class kernels +
digestion + symmetrisation) of the seeded (H2O)64 cluster at the converged
RHF density, with the density and J/K resident in HBM.  Before the timed
region the RHF is converged in FP64 (plugin defaults: convergence 1e-7,
eri_threshold 1e-12, shell_pair_threshold 1e-12); with mixed precision the
same SCF is repeated with FP32 for quartets whose Schwarz bound is below 1e-6
and the energy difference is reported.  The same builds are also timed at the
atomic-guess density, and one ERI nuclear gradient (the adjoint kernel) at
the converged density.  The builds are timed twice: on an engine with
per-class HIP events (its class kernels run one after another: per-class
times, FLOPs and the roofline) and on a plain engine (its class kernels run
concurrently on the engine's streams: the headline value).  Both use the
default use_atomics = False (bitwise-repeatable fixed-point sums, the
reference's default); the plain FP64 builds are repeated with use_atomics =
True.  The CPU baseline is the oracle (a C OpenMP
restatement of the reference's LIBINT2_DIRECT loop with its own Obara-Saika
integrals, not Libint2) on a strided sample of bra shells of the same builds,
at both densities, in a child process with the box's CPU share.  Prints one
JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "qdk-chemistry_amd"))

FP64_PEAK = 78.6e12  # MI355X dense FP64 vector, FLOP/s
FP32_PEAK = 157.3e12  # MI355X FP32 vector, FLOP/s
METRIC = "Fock-build s and ERI quartets/s, (H2O)64/cc-pVDZ, mixed vs FP64 at 1 µEh"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def system(name):
    from qdk_chemistry_amd.molecule import water_cluster
    if name == "h2o64":
        return water_cluster(64, 20261264), "cc-pvdz", "(H2O)64/cc-pVDZ seed 20261264"
    if name == "h2o16":
        return water_cluster(16, 20261202), "cc-pvdz", "(H2O)16/cc-pVDZ seed 20261202"
    if name == "c20h42":
        from qdk_chemistry_amd.molecule import linear_alkane
        return linear_alkane(20, 20261220), "cc-pvtz", "C20H42/cc-pVTZ seed 20261220"
    raise SystemExit(f"unknown system {name}")


def timed_builds(eng, P, steps, warmup):
    """K builds on HBM-resident P/J/K; per-class counts and kernel times
    (HIP events recorded on the engine stream) for both precisions."""
    # The engine's own HBM buffers and stream: torch's bundled HIP runtime is a
    # different build from the system ROCm the library links, so device
    # memory and synchronisation stay inside the library.  build_JK_device
    # returns after the build's final event has completed (device synchronous).
    dP, dJ, dK = eng.device_buffers()
    eng.upload(dP, P)
    for _ in range(warmup):
        eng.build_JK_device(dP, dJ, dK)
    per_class = {}
    quartets = quartets32 = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.build_JK_device(dP, dJ, dK)
        for prec in (0, 1):
            for cls, nq, npq, ms in eng.class_stats(prec):
                if nq > 0:
                    acc = per_class.setdefault((prec, cls), [0.0, 0.0, 0.0, 0])
                    acc[0] += nq
                    acc[2] += ms
                    acc[3] += 1
        st = eng.stats()
        quartets += st["quartets"]
        quartets32 += st["quartets_fp32"]
    dt = time.perf_counter() - t0
    # the evaluated primitive quartets (the same in every build of one
    # density), counted after the timed region
    for prec in (0, 1):
        for cls, nq, npq, ms in eng.class_stats(prec, executed=True):
            if (prec, cls) in per_class:
                per_class[(prec, cls)][1] = npq * per_class[(prec, cls)][3]
    n = P.shape[0]
    return dict(seconds=dt, quartets=quartets, quartets_fp32=quartets32, per_class=per_class, stats=eng.stats(),
                J=eng.download(dJ, (n, n)), K=eng.download(dK, (n, n)))


def wall_builds(eng, P, steps, warmup):
    """K builds on HBM-resident P/J/K on an engine without per-class events
    (its class kernels then run concurrently on the engine's streams):
    ms per build and quartets/s."""
    dP, dJ, dK = eng.device_buffers()
    eng.upload(dP, P)
    for _ in range(warmup):
        eng.build_JK_device(dP, dJ, dK)
    quartets = quartets32 = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.build_JK_device(dP, dJ, dK)
        st = eng.stats()
        quartets += st["quartets"]
        quartets32 += st["quartets_fp32"]
    dt = time.perf_counter() - t0
    n = P.shape[0]
    return dict(ms_per_build=1e3 * dt / steps, quartets_per_s=quartets / dt, quartets_per_build=quartets / steps,
                fp32_fraction=quartets32 / max(quartets, 1), J=eng.download(dJ, (n, n)), K=eng.download(dK, (n, n)))


def summarize(run, steps):
    from qdk_chemistry_amd.flops import class_flops
    f64 = f32 = 0.0
    dom = None
    dom_total = -1.0
    classes = {}
    for (prec, cls), (nq, npq, ms, n) in run["per_class"].items():
        fp, fc = class_flops(*cls)
        f = npq * fp + nq * fc
        classes["(%d%d|%d%d) %s" % (*cls, "fp32" if prec else "fp64")] = dict(
            quartets=nq / steps, prim_quartets=npq / steps, ms=ms / steps, flops=f / steps)
        if prec == 0:
            f64 += f
            if ms > dom_total:  # the FP64 class kernel group with the most time
                dom_total = ms
                dom = (cls, f / n, nq / n, ms / n)
        else:
            f32 += f
    eri_ms = sum(v[2] for v in run["per_class"].values()) / steps
    ms_step = 1e3 * run["seconds"] / steps
    t_peak = (f64 / FP64_PEAK + f32 / FP32_PEAK) / steps
    return dict(ms_per_build=ms_step, quartets_per_s=run["quartets"] / run["seconds"],
                quartets_per_build=run["quartets"] / steps, fp32_fraction=run["quartets_fp32"] / max(run["quartets"], 1),
                eri_kernels_ms=eri_ms, screen_ms=run["stats"]["ms_screen"],
                flops_fp64_per_build=f64 / steps, flops_fp32_per_build=f32 / steps,
                peak_fraction_eri=t_peak / (eri_ms * 1e-3) if eri_ms > 0 else 0.0,
                classes=classes, dominant=dom)


def calibration_copy(eng, nbf):
    """One plain device copy of the density buffer into the J buffer (k_copy_words,
    8*nbf^2 bytes each way): the reference reading for HBM traffic counters."""
    from qdk_chemistry_amd import _lib
    dP, dJ, _ = eng.device_buffers()
    nbytes = 8 * nbf * nbf
    _lib.check(_lib.lib().qfe_copy(eng.h, dJ, dP, nbytes, 2))
    return nbytes


def traffic_from_profile(kernel_tag):
    """HBM bytes per launch of the dominant kernel, from the committed PMC
    summary (profiles/hbm_traffic.txt: 'tag = bytes' lines written from
    rocprofv3 FETCH_SIZE and WRITE_SIZE runs scaled by the copy-kernel reading)."""
    path = os.path.join(REPO, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            if k.strip() == kernel_tag:
                return float(v.split()[0])
    return None


def log_classes(run, steps, top=12):
    """Per-class table on stderr: quartets, ms and TFLOP/s per build, and the
    exchange updates the digestion adds into the LDS rows of its segment
    (four K blocks per shell quartet: (na + nb) x (nc + nd) LDS atomics; the
    global atomics are the flushes of those rows, once per non-zero entry and
    segment, the J_ab block sums and the J_cd tile sums)."""
    from qdk_chemistry_amd.flops import class_flops, nsph
    rows = []
    for (prec, cls), (nq, npq, ms, n) in run["per_class"].items():
        fp, fc = class_flops(*cls)
        na, nb, nc, nd = (nsph(x) for x in cls)
        upd = nq * (na + nb) * (nc + nd)
        rows.append((ms / steps, prec, cls, nq / steps, npq / steps, (npq * fp + nq * fc) / (ms * 1e-3) / 1e12,
                     upd / steps))
    rows.sort(reverse=True)
    tot = sum(r[6] for r in rows)
    log(f"classes by time (of {len(rows)}); exchange LDS updates per build {tot:.3g}")
    for ms, prec, cls, nq, npq, tf, up in rows[:top]:
        log(f"  {'f32' if prec else 'f64'} ({cls[0]}{cls[1]}|{cls[2]}{cls[3]}) {ms:8.2f} ms  quartets {nq:.3g}  "
            f"prim {npq:.3g}  {tf:6.2f} TFLOP/s  LDS updates {up:.3g}")


def harness_timings(basis, mol, P, host=True):
    """Seconds for the SCF harness pieces that run beside the Fock build on
    the GPU: S + T + V, their three gradient contractions, and one
    diagonalisation of an nbf x nbf symmetric matrix; with their host
    counterparts (16-thread C++ integrals, numpy/LAPACK eigh) when host."""
    import numpy as np
    from qdk_chemistry_amd.eri import eigh_gpu, one_body, one_body_gpu, one_body_grad_gpu

    def timed(f):
        f()  # first call: tables, handles
        t = time.perf_counter()
        f()
        return time.perf_counter() - t

    A = 0.5 * (P + P.T)
    out = dict(one_body_gpu_s=timed(lambda: [one_body_gpu(basis, mol, op) for op in range(3)]),
               one_body_grad_gpu_s=timed(lambda: [one_body_grad_gpu(basis, mol, op, P) for op in range(3)]),
               eigh_gpu_s=timed(lambda: eigh_gpu(A)))
    if host:
        out.update(one_body_host_s=timed(lambda: [one_body(basis, mol, op) for op in range(3)]),
                   eigh_host_s=timed(lambda: np.linalg.eigh(A)))
    log("harness: " + ", ".join(f"{k} {v:.3f}" for k, v in out.items()))
    return out


def mo_transform_timings(nwat, nt, host=True):
    """Seconds for the first-quarter transform (nbf^3 x nt) and the
    four-index MO transform (nt^4) of a water cluster in cc-pVDZ on the GPU
    (dense AO tensor in HBM + rocBLAS GEMMs, output copied to the host),
    and, when host, the oracle's full tensor plus numpy contractions on the
    host cores for the same transform (a CPU baseline)."""
    import numpy as np
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import MOERI, create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(nwat, 4000 + nwat)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    nt = min(nt, b.nbf)
    C = np.random.Generator(np.random.PCG64(nt)).uniform(-1.0, 1.0, (b.nbf, nt))
    eri = create(b, mol)
    mo = MOERI(eri)
    eri.quarter_trans(C[:, :1])  # tables and handles
    t = time.perf_counter()
    q = eri.quarter_trans(C)
    tq = time.perf_counter() - t
    t = time.perf_counter()
    m = mo.compute(b.nbf, nt, C, C, C, C)
    tm = time.perf_counter() - t
    out = dict(workload=f"(H2O){nwat} cc-pVDZ, nbf {b.nbf}, nt {nt}", quarter_trans_gpu_s=tq, mo_eri_gpu_s=tm)
    if host:
        from oracle.oracle import OracleERI, mo_eri
        threads = min(16, os.cpu_count() or 1)
        t = time.perf_counter()
        T = OracleERI(b, mol, nthreads=threads).eri_full()
        ref = mo_eri(T, C, C, C, C)
        out["mo_eri_host_s"] = time.perf_counter() - t
        out["host_threads"] = threads
        out["max_abs_diff_vs_host"] = float(np.max(np.abs(ref - m)))
        del T, ref
    del q, m
    log("mo transform: " + ", ".join(f"{k} {v:.3g}" if isinstance(v, float) else f"{k} {v}"
                                      for k, v in out.items()))
    return out


def dfj_timings(nwat):
    """Seconds for density-fitted J on a water cluster in def2-SVP with
    def2-universal-jfit: the set-up (three-centre integrals, metric and its
    Cholesky factor, once) and one J build, beside one exact J/K build of the
    Rys engine at the same density; and the fitting error of J."""
    import numpy as np
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import DFJ, create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(nwat, 5000 + nwat)
    b = load_basis("def2-svp", mol.numbers, mol.coords)
    aux = load_basis("def2-universal-jfit", mol.numbers, mol.coords)
    rng = np.random.Generator(np.random.PCG64(nwat))
    A = rng.uniform(-0.05, 0.05, (b.nbf, b.nbf))
    P = A + A.T
    t = time.perf_counter()
    df = DFJ(b, aux, mol)
    t_setup = time.perf_counter() - t
    df.build_JK(P)
    t = time.perf_counter()
    Jd, _ = df.build_JK(P)
    t_j = time.perf_counter() - t
    eri = create(b, mol)
    eri.build_JK(P)
    t = time.perf_counter()
    Je, _ = eri.build_JK(P)
    t_jk = time.perf_counter() - t
    out = dict(workload=f"(H2O){nwat} def2-SVP / def2-universal-jfit, nbf {b.nbf}, naux {df.naux}",
               setup_s=t_setup, build_j_s=t_j, exact_jk_build_s=t_jk,
               max_rel_j_error=float(np.max(np.abs(Jd - Je)) / np.max(np.abs(Je))))
    log("DF-J: " + ", ".join(f"{k} {v:.3g}" if isinstance(v, float) else f"{k} {v}" for k, v in out.items()))
    return out


def c20h42_scf(mix_threshold=1e-6):
    """C20H42/cc-pVTZ (f shells): the FP64 RHF and the mixed-precision RHF
    (FP32 below Q_ab Q_cd = mix_threshold; the f-shell classes evaluate their
    FP32 tasks in FP64) from the atomic guess, with the energy difference."""
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig, create
    from qdk_chemistry_amd.scf import RHF, SCFConfig
    import numpy as np
    mol, bn, desc = system("c20h42")
    b = load_basis(bn, mol.numbers, mol.coords)
    cfg = ERIConfig(eri_threshold=1e-12, shell_pair_threshold=1e-12)
    out = dict(workload=desc, nbf=b.nbf)
    for label, mix in (("fp64", False), ("mixed", True)):
        eng = create(b, mol, cfg)
        if mix:
            eng.set_precision(True, mix_threshold)
        t = time.perf_counter()
        scf = RHF(mol, b, SCFConfig(eri=cfg), eri=eng)
        r = scf.run()
        out[label] = dict(energy=r["energy"], iterations=r["iterations"], wall_s=time.perf_counter() - t,
                          fock_build_s=float(np.sum(scf.build_times)),
                          mixed_builds=r.get("mixed_builds", 0))
        eng.close()
    out["dE_mixed_vs_fp64"] = out["mixed"]["energy"] - out["fp64"]["energy"]
    out["fock_speedup_mixed"] = out["fp64"]["fock_build_s"] / out["mixed"]["fock_build_s"]
    log(f"C20H42/cc-pVTZ SCF: FP64 E={out['fp64']['energy']:.10f} (Fock {out['fp64']['fock_build_s']:.2f}s), "
        f"mixed dE={out['dE_mixed_vs_fp64']:+.2e} (Fock {out['mixed']['fock_build_s']:.2f}s, "
        f"{out['mixed']['mixed_builds']} mixed builds)")
    return out


def gradient_timings(systems=("h2o16", "c20h42")):
    """The ERI gradient (adjoint kernels) of the gradient configurations,
    (H2O)16/cc-pVDZ and C20H42/cc-pVTZ (f shells: the generic kernel), at
    the atomic-guess density: one warm-up, then one timed call each."""
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig, create
    from qdk_chemistry_amd.scf import atom_guess
    out = {}
    for name in systems:
        mol, bn, desc = system(name)
        b = load_basis(bn, mol.numbers, mol.coords)
        eng = create(b, mol, ERIConfig(eri_threshold=1e-12, shell_pair_threshold=1e-12))
        P = atom_guess(mol, bn)
        eng.get_gradients(P)
        t = time.perf_counter()
        eng.get_gradients(P)
        dt = time.perf_counter() - t
        st = eng.stats()
        eng.close()
        out[name] = dict(workload=desc, nbf=b.nbf, density="atomic guess", seconds=dt, quartets=st["quartets"],
                         quartets_per_s=st["quartets"] / dt, ms_kernels=st["ms_eri"])
        log(f"ERI gradient {desc}: {dt * 1e3:.1f} ms, {st['quartets']:.4g} quartets, kernels {st['ms_eri']:.1f} ms")
    return out


def cpu_baseline(system_name, densities, stride, repeats=5):
    """The oracle's build_JK on a bounded sample of the same build, run in a
    child process (oracle/cpu_baseline.py) so that OMP_NUM_THREADS and
    OMP_PROC_BIND=close reach its OpenMP runtime: the box's CPU share (the
    OMP_NUM_THREADS it sets, else every core), one warm-up and the median of
    `repeats` timed builds per density."""
    import subprocess
    import tempfile
    import numpy as np
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    with tempfile.TemporaryDirectory() as tmp:
        args = [sys.executable, os.path.join(REPO, "oracle", "cpu_baseline.py"), "--system", system_name,
                "--stride", str(stride), "--repeats", str(repeats)]
        for label, P in densities.items():
            path = os.path.join(tmp, f"{label}.npy")
            np.save(path, P)
            args += ["--density", f"{label}={path}"]
        r = subprocess.run(args, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"CPU baseline failed: {r.stderr[-2000:]}")
    res = json.loads(r.stdout.strip().splitlines()[-1])
    head = res["densities"]["converged"] if "converged" in res["densities"] else next(iter(res["densities"].values()))
    dens = "; ".join(f"{k}: {v['quartets']:.0f} quartets, median {v['median_s']:.2f} s, {v['quartets_per_s']:.4g}/s"
                     for k, v in res["densities"].items())
    return dict(value=head["quartets_per_s"], unit="quartets/s", cores=res["cores"], kind="port",
                cpu_model=res["cpu_model"], omp_proc_bind=res["omp_proc_bind"],
                sample=f"oracle build_JK (C/OpenMP restatement of LIBINT2_DIRECT, Obara-Saika integrals, "
                       f"not Libint2) over the bra shells s1 % {stride} == 0 of the same build, 1 warm-up + "
                       f"median of {repeats}, {res['cores']} threads on {res['cpu_model']}: {dens}",
                densities=res["densities"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--system", default="h2o64")
    ap.add_argument("--no-scf", action="store_true", help="time at the atomic-guess density")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-grad", action="store_true", help="skip the ERI gradient timing")
    ap.add_argument("--no-atomics", action="store_true", help="skip the use_atomics = True builds")
    ap.add_argument("--no-c5", action="store_true", help="skip the C20H42/cc-pVTZ FP64 and mixed SCF")
    ap.add_argument("--profiled-only", action="store_true",
                    help="time only the profiled engine (one stream; for profiler runs: one run of each class kernel "
                         "per build)")
    ap.add_argument("--cpu-stride", type=int, default=96)
    ap.add_argument("--mo-waters", type=int, default=4, help="water cluster of the MO-transform timing (0: skip)")
    ap.add_argument("--mo-nt", type=int, default=48)
    ap.add_argument("--mo-only", action="store_true", help="only the MO-transform timing, one JSON line")
    ap.add_argument("--dfj-waters", type=int, default=16, help="water cluster of the DF-J timing (0: skip)")
    ap.add_argument("--dfj-only", action="store_true", help="only the DF-J timing, one JSON line")
    ap.add_argument("--density", help="load the density from this .npy instead of running the SCF")
    ap.add_argument("--save-density", help="write the converged density to this .npy")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path is single-device "
              f"in this build.  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.dfj_only:
        print(json.dumps({"dfj": dfj_timings(args.dfj_waters)}), flush=True)
        return
    if args.mo_only:
        print(json.dumps({"mo_transform": mo_transform_timings(args.mo_waters, args.mo_nt, not args.no_cpu)}),
              flush=True)
        return

    import numpy as np
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig, create
    from qdk_chemistry_amd.flops import class_flops
    from qdk_chemistry_amd.scf import RHF, SCFConfig, atom_guess

    mol, basis_name, desc = system(args.system)
    basis = load_basis(basis_name, mol.numbers, mol.coords)
    cfg = ERIConfig(eri_threshold=1e-12, shell_pair_threshold=1e-12, profile=True)
    t = time.perf_counter()
    eng = create(basis, mol, cfg)
    log(f"engine: {desc}, nbf={basis.nbf}, setup {time.perf_counter() - t:.2f}s, "
        f"pairs={eng.stats()['significant_pairs']}")

    result = {}
    if args.density:
        P = np.load(args.density)
        if P.shape != (basis.nbf, basis.nbf):
            raise SystemExit(f"{args.density}: shape {P.shape} does not match nbf={basis.nbf}")
        density = f"density loaded from {os.path.basename(args.density)}"
    elif args.no_scf:
        P = atom_guess(mol, basis_name)
        density = "atomic guess"
    else:
        t = time.perf_counter()
        scf = RHF(mol, basis, SCFConfig(eri=cfg), eri=eng)
        r = scf.run()
        P = r["P"]
        density = "converged FP64 RHF"
        result["scf_fp64"] = dict(energy=r["energy"], iterations=r["iterations"],
                                  fock_build_s=float(np.sum(scf.build_times)), wall_s=time.perf_counter() - t)
        log(f"FP64 SCF: E={r['energy']:.10f} in {r['iterations']} iterations, "
            f"Fock builds {np.sum(scf.build_times):.2f}s")
        if args.save_density:
            np.save(args.save_density, P)

    run64 = timed_builds(eng, P, args.steps, args.warmup)
    s64 = summarize(run64, args.steps)
    calib_bytes = calibration_copy(eng, basis.nbf)
    cls, f_launch, nq_launch, ms_launch = s64["dominant"]
    achieved = f_launch / (ms_launch * 1e-3)
    log(f"FP64 timed: {args.steps} builds, {s64['ms_per_build']:.2f} ms/build, {s64['quartets_per_s']:.4g} quartets/s, "
        f"ERI kernels {s64['eri_kernels_ms']:.2f} ms ({s64['peak_fraction_eri']:.3f} of peak), screen "
        f"{s64['screen_ms']:.2f} ms; dominant class {cls}: {ms_launch:.2f} ms, {achieved / 1e12:.2f} TFLOP/s")
    log_classes(run64, args.steps)
    mixed = None
    if not args.no_mixed:
        eng.set_precision(True, 1e-6)
        run32 = timed_builds(eng, P, args.steps, args.warmup)
        eng.set_precision(False)
        s32 = summarize(run32, args.steps)
        import numpy as np
        Jm, Km = run32["J"], run32["K"]
        J6, K6 = run64["J"], run64["K"]
        mixed = dict(s32)
        mixed.pop("dominant")
        mixed["max_abs_dF_vs_fp64"] = float(np.max(np.abs((Jm - 0.5 * Km) - (J6 - 0.5 * K6))))
        mixed["speedup_vs_fp64"] = s64["ms_per_build"] / s32["ms_per_build"]
        log(f"mixed timed: {s32['ms_per_build']:.2f} ms/build ({mixed['speedup_vs_fp64']:.2f}x), "
            f"FP32 share {s32['fp32_fraction']:.3f}, max|dF| {mixed['max_abs_dF_vs_fp64']:.2e}")
        log_classes(run32, args.steps)
        if "scf_fp64" in result:  # the mixed SCF is compared with the FP64 one run above
            t = time.perf_counter()
            eng.set_precision(True, 1e-6)
            scf_m = RHF(mol, basis, SCFConfig(eri=cfg), eri=eng)
            rm = scf_m.run()
            eng.set_precision(False)
            dE = rm["energy"] - result["scf_fp64"]["energy"]
            mixed["scf"] = dict(energy=rm["energy"], dE_vs_fp64=dE, iterations=rm["iterations"],
                                mixed_builds=rm["mixed_builds"], fock_build_s=float(np.sum(scf_m.build_times)),
                                wall_s=time.perf_counter() - t)
            log(f"mixed SCF: E={rm['energy']:.10f} dE={dE:+.2e} ({rm['iterations']} iterations, "
                f"{rm['mixed_builds']} mixed builds, Fock builds {np.sum(scf_m.build_times):.2f}s)")
    # the same builds at the atomic-guess density (the first SCF iteration's
    # full density; the screening keeps more quartets there)
    guess = None
    if "scf_fp64" in result:
        Pg = atom_guess(mol, basis_name)
        rg = timed_builds(eng, Pg, args.steps, args.warmup)
        sg = summarize(rg, args.steps)
        guess = dict(fp64={k: v for k, v in sg.items() if k != "dominant"})
        log(f"atomic-guess density, FP64: {sg['ms_per_build']:.2f} ms/build, {sg['quartets_per_s']:.4g} quartets/s "
            f"({sg['quartets_per_build']:.4g} quartets)")
        if not args.no_mixed:
            eng.set_precision(True, 1e-6)
            rgm = timed_builds(eng, Pg, args.steps, args.warmup)
            eng.set_precision(False)
            sgm = summarize(rgm, args.steps)
            guess["mixed"] = {k: v for k, v in sgm.items() if k != "dominant"}
            guess["mixed"]["speedup_vs_fp64"] = sg["ms_per_build"] / sgm["ms_per_build"]
            guess["mixed"]["max_abs_dF_vs_fp64"] = float(np.max(np.abs((rgm["J"] - 0.5 * rgm["K"]) -
                                                                       (rg["J"] - 0.5 * rg["K"]))))
            log(f"atomic-guess density, mixed: {sgm['ms_per_build']:.2f} ms/build "
                f"({guess['mixed']['speedup_vs_fp64']:.2f}x)")
    # ERI nuclear gradient (the adjoint kernel) at the same density
    grad = None
    if not args.no_grad:
        eng.get_gradients(P)
        t = time.perf_counter()
        eng.get_gradients(P)
        gst = eng.stats()
        grad = dict(seconds=time.perf_counter() - t, quartets=gst["quartets"], ms_screen=gst["ms_screen"],
                    ms_kernels=gst["ms_eri"])
        log(f"ERI gradient: {grad['seconds'] * 1e3:.1f} ms ({gst['quartets']:.4g} quartets, "
            f"kernels {gst['ms_eri']:.1f} ms)")
    eng.close()
    if args.profiled_only:
        out = {"metric": METRIC, "value": s64["quartets_per_s"], "unit": "quartets/s", "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": s64["ms_per_build"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
               "data": f"synthetic seeded geometry, {density} density",
               "config": {"workload": f"RHF J/K build, {desc}, nbf={basis.nbf}, eri_threshold 1e-12",
                          "path": "fp64 profiled engine",
                          "fp64_profiled": {k: v for k, v in s64.items() if k != "dominant"}},
               "roofline": {"bound": "fp64-valu", "kernel": "ERI class (%d %d|%d %d), FP64" % cls,
                            "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                            "frac": achieved / FP64_PEAK, "flops_per_launch": f_launch, "ms_per_launch": ms_launch}}
        print(json.dumps(out), flush=True)
        return
    # The headline: the same builds on an engine without per-class events, so
    # that its class kernels run concurrently on the engine's streams (the
    # profiled engine above launches them one after another, for clean
    # per-class times and the roofline)
    eng_p = create(basis, mol, ERIConfig(eri_threshold=1e-12, shell_pair_threshold=1e-12))
    w64 = wall_builds(eng_p, P, args.steps, args.warmup)
    prod = dict(fp64={k: v for k, v in w64.items() if k not in ("J", "K")})
    log(f"FP64 builds, concurrent classes: {w64['ms_per_build']:.2f} ms/build, {w64['quartets_per_s']:.4g} quartets/s "
        f"(max|dJ|, max|dK| vs the profiled builds {np.max(np.abs(w64['J'] - run64['J'])):.1e}, "
        f"{np.max(np.abs(w64['K'] - run64['K'])):.1e})")
    headline, path = w64, "fp64"
    if not args.no_mixed:
        eng_p.set_precision(True, 1e-6)
        w32 = wall_builds(eng_p, P, args.steps, args.warmup)
        eng_p.set_precision(False)
        prod["mixed"] = {k: v for k, v in w32.items() if k not in ("J", "K")}
        prod["mixed"]["speedup_vs_fp64"] = w64["ms_per_build"] / w32["ms_per_build"]
        log(f"mixed builds, concurrent classes: {w32['ms_per_build']:.2f} ms/build "
            f"({prod['mixed']['speedup_vs_fp64']:.2f}x)")
        if (mixed is None or "scf" not in mixed or abs(mixed["scf"]["dE_vs_fp64"]) <= 1e-6) and \
                w32["ms_per_build"] < w64["ms_per_build"]:
            headline, path = w32, "mixed"
    eng_p.close()
    # the same FP64 builds with use_atomics = True (double-precision atomics:
    # no fixed-point conversion, not bitwise repeatable), on a fresh engine
    atomics = None
    if not args.no_atomics:
        eng_a = create(basis, mol, ERIConfig(eri_threshold=1e-12, shell_pair_threshold=1e-12, use_atomics=True))
        wa = wall_builds(eng_a, P, args.steps, args.warmup)
        eng_a.close()
        atomics = {k: v for k, v in wa.items() if k not in ("J", "K")}
        log(f"use_atomics=True, FP64, concurrent classes: {wa['ms_per_build']:.2f} ms/build, "
            f"{wa['quartets_per_s']:.4g} quartets/s")
    out = {
        "metric": METRIC,
        "value": headline["quartets_per_s"],
        "unit": "quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": headline["ms_per_build"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if path == "fp64" else "f64+f32",
        "data": f"synthetic seeded geometry, {density} density",
        "config": {"workload": f"RHF J/K build, {desc}, nbf={basis.nbf}, eri_threshold 1e-12",
                   "path": path, "fock_build_s": headline["ms_per_build"] / 1e3, "builds": prod,
                   "fp64_profiled": {k: v for k, v in s64.items() if k != "dominant"}, "mixed_profiled": mixed,
                   "atom_guess_density": guess, "eri_gradient": grad, "fp64_use_atomics": atomics},
        "roofline": {"bound": "fp64-valu", "kernel": "ERI class (%d %d|%d %d), FP64" % cls,
                     "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK,
                     "traffic": traffic_from_profile("%s (%d%d|%d%d) fp64" % (args.system, *cls)),
                     "flops_per_launch": f_launch, "ms_per_launch": ms_launch,
                     "calibration_copy_bytes": calib_bytes},
    }
    out["config"].update(result)
    if "scf_fp64" in result:
        out["config"]["harness"] = harness_timings(basis, mol, P, not args.no_cpu)
    if not args.no_grad:
        out["config"]["eri_gradient_configs"] = gradient_timings()
    if not args.no_c5:
        out["config"]["c20h42_scf"] = c20h42_scf()
    if args.dfj_waters > 0:
        out["config"]["dfj"] = dfj_timings(args.dfj_waters)
    if args.mo_waters > 0:
        out["config"]["mo_transform"] = mo_transform_timings(args.mo_waters, args.mo_nt, not args.no_cpu)
    if not args.no_cpu:
        dens = {"converged": P} if "scf_fp64" in result else {"given": P}
        if "scf_fp64" in result:
            dens["atom_guess"] = Pg
        cb = cpu_baseline(args.system, dens, args.cpu_stride)
        out["cpu_baseline"] = cb
        log(f"cpu baseline: {cb['value']:.4g} quartets/s on {cb['cores']} threads ({cb['cpu_model']})")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
