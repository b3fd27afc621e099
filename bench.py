#!/usr/bin/env python3
"""Benchmark of the RI hot path on one MI355X.

Workload (default): the RI-K build at nbas=1024, naux=4096, nocc=128 with
ri3ao (34.4 GB, FP64) and the occupied-row ri3mo B (4.3 GB) resident in HBM.
One step = one K = sum_P B_P^T B_P build (FP64 MFMA), credited with
F_K = nbas(nbas+1)*nocc*naux flops (upper triangle only).  The ao2mo, RI-J
and pipeline rates are reported beside it, and the oracle's CPU K build is
timed on a bounded sample of the same workload as the CPU baseline.

Prints ONE JSON line.  Only one GPU is supported: --gpus N > 1 exits.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 MFMA
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the FP64 RI-K kernel, measured with rocprofv3 --pmc
# (profiles/ri_pipeline_nbas1024_naux4096_nocc128.md): FETCH_SIZE 5181450 KiB
# (mean of two dispatches) x 2.000 (a plain 512 MiB copy reads as 262182 KiB)
# + WRITE_SIZE 278528 KiB x 1.000 (exact on the same copy).  Keyed by
# (nbas, naux, nocc); other shapes report null.
MEASURED_TRAFFIC_RIK = {(1024, 4096, 128): (5181450 * 2.000 + 278528 * 1.000) * 1024}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nbas", type=int, default=1024)
    ap.add_argument("--naux", type=int, default=4096)
    ap.add_argument("--nocc", type=int, default=128)
    ap.add_argument("--seed", type=int, default=0x5EED0003)
    ap.add_argument("--cpu-sample", type=int, default=4096,
                    help="aux slices of the CPU-baseline K sample (0 = skip)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the ao2mo / RI-J side measurements")
    ap.add_argument("--batch-mols", type=int, default=64,
                    help="molecules of the batched workload (0 = skip)")
    ap.add_argument("--davidson", type=int, default=1,
                    help="run the Davidson workload (plain vs learned preconditioner)")
    ap.add_argument("--forms", default="ozaki_int8,split_fp32,ozaki_bf16,bf16x3",
                    help="comma-separated reduced-precision K forms to time ('' = none)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run one plain device copy of 512 MiB (copy3d kernel: reads "
                         "512 MiB, writes 512 MiB) to calibrate FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()

    if args.traffic_probe:
        from rest_tensors_amd import device as D
        from rest_tensors_amd._lib import lib, check
        n = 64 * 1024 * 1024
        src, dst = D.DeviceArray(n), D.DeviceArray(n)
        D.gen_normal(src, n, seed=1)
        D.synchronize()
        check(lib().rt_copy_mm_dev(n, 1, src.ptr, n, 1, 0, 0, dst.ptr, n, 1, 0, 0, None))
        D.synchronize()
        print(json.dumps({"traffic_probe": "copy3d_kernel", "read_KiB": n * 8 // 1024,
                          "write_KiB": n * 8 // 1024}))
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; "
                          "multi-GPU sharding over the aux index is designed, not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from rest_tensors_amd import device as D
    from oracle import ri_oracle as O   # only for C (host QR) and the CPU baseline

    nbas, naux, nocc = args.nbas, args.naux, args.nocc
    nmo = nbas
    log(f"[bench] device {D.device_name()}  nbas={nbas} naux={naux} nocc={nocc}")

    # ---- inputs resident in HBM -----------------------------------------------
    t0 = time.time()
    dA = D.DeviceArray(nbas * nbas * naux)
    D.gen_ri3ao(dA, nbas, naux, seed=args.seed)
    Cm = O.gen_orthonormal(nbas, seed=args.seed + 1)
    dC = D.DeviceArray.from_host(Cm)
    dB = D.DeviceArray(nocc * nmo * naux)
    dK = D.DeviceArray(nmo * nmo)
    D.ao2mo_occ(dC, nbas, nmo, nocc, dA, naux, dB)
    D.synchronize()
    log(f"[bench] inputs generated in {time.time() - t0:.1f}s")

    f_k = float(nmo) * (nmo + 1) * nocc * naux
    f_ao2mo = 4.0 * nocc * nbas * nbas * naux
    bytes_j = 8.0 * naux * nbas * (nbas + 1)

    # ---- timed RI-K steps ---------------------------------------------------------
    for _ in range(args.warmup):
        D.rik(dB, nocc, nmo, naux, dK, want_energy=False)
    D.synchronize()
    D.kernel_timer_reset()
    D.kernel_timer(True)
    ev0, ev1 = D.Event(), D.Event()
    D.synchronize()
    ev0.record()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        D.rik(dB, nocc, nmo, naux, dK, want_energy=False)
    ev1.record()
    D.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev0.elapsed_ms(ev1)
    D.kernel_timer(False)
    k_ms, k_cnt = D.kernel_time("rik_mfma")
    r_ms, r_cnt = D.kernel_time("reduce_sym")
    D.kernel_timer_reset()
    ms_per_step = ms_total / args.steps
    value = f_k / (ms_per_step * 1e-3) / 1e12
    k_avg_ms = k_ms / max(k_cnt, 1)
    achieved = f_k / (k_avg_ms * 1e-3) / 1e12
    log(f"[bench] RI-K: {ms_per_step:.3f} ms/step ({value:.2f} TFLOP/s), "
        f"rik_mfma avg {k_avg_ms:.3f} ms, reduce avg {r_ms / max(r_cnt, 1):.3f} ms, "
        f"wall {wall:.2f}s")

    e_x = D.rik(dB, nocc, nmo, naux, dK, want_energy=True)
    extras = {"e_x": e_x}

    # ---- reduced-precision K builds (slicing pass included in the step) -----------
    K64 = dK.to_host((nmo, nmo))
    prec = {}
    dK2 = D.DeviceArray(nmo * nmo)
    knames = {"ozaki_int8": "rik_oz_i8", "split_fp32": "rik_split32",
              "ozaki_bf16": "rik_ozaki", "bf16x3": "rik_bf16x3"}
    for form in [f for f in args.forms.split(",") if f]:
        kname = knames[form]
        for _ in range(max(1, args.warmup)):
            D.rik(dB, nocc, nmo, naux, dK2, precision=form, want_energy=False)
        D.synchronize()
        D.kernel_timer_reset()
        D.kernel_timer(True)
        ev0.record()
        for _ in range(args.steps):
            D.rik(dB, nocc, nmo, naux, dK2, precision=form, want_energy=False)
        ev1.record()
        ms = ev0.elapsed_ms(ev1) / args.steps
        D.kernel_timer(False)
        kk_ms, kk_n = D.kernel_time(kname)
        sl_ms, sl_n = D.kernel_time("rik_slice")
        D.kernel_timer_reset()
        ex2 = D.rik(dB, nocc, nmo, naux, dK2, precision=form, want_energy=True)
        K2 = dK2.to_host((nmo, nmo))
        prec[form] = {
            "ms_per_step": round(ms, 4),
            "tflops_fp64_equiv": round(f_k / (ms * 1e-3) / 1e12, 3),
            "speedup_vs_fp64": round(ms_per_step / ms, 3),
            "kernel_avg_ms": round(kk_ms / max(kk_n, 1), 4),
            "slice_avg_ms": round(sl_ms / max(sl_n, 1), 4),
            "K_rel_frobenius_err": float(np.linalg.norm(K2 - K64) / np.linalg.norm(K64)),
            "E_x_rel_err": abs(ex2 - e_x) / abs(e_x),
        }
        log(f"[bench] RI-K {form}: {ms:.3f} ms/step ({prec[form]['tflops_fp64_equiv']} "
            f"TFLOP/s FP64-equiv, x{prec[form]['speedup_vs_fp64']} vs FP64), "
            f"K err {prec[form]['K_rel_frobenius_err']:.2e}")
    extras["precision_forms"] = prec
    del dK2

    # ---- side measurements: ao2mo, RI-J, unpack ---------------------------------------
    if not args.no_extras:
        D.kernel_timer(True)
        ev0.record()
        D.ao2mo_occ(dC, nbas, nmo, nocc, dA, naux, dB)
        ev1.record()
        ao_ms = ev0.elapsed_ms(ev1)
        dD = D.DeviceArray(nbas * nbas)
        dJ = D.DeviceArray(nbas * nbas)
        D.density(dC, nbas, nocc, dD)
        D.rij(dA, nbas, naux, dD, dJ, want_energy=False)   # warm
        D.synchronize()
        D.kernel_timer_reset()
        ev0.record()
        e_j = D.rij(dA, nbas, naux, dD, dJ, want_energy=True)
        ev1.record()
        j_ms = ev0.elapsed_ms(ev1)
        g_ms, _ = D.kernel_time("rij_gamma")
        jj_ms, _ = D.kernel_time("rij_j")
        D.kernel_timer(False)
        D.kernel_timer_reset()
        extras.update({
            "ao2mo_ms": round(ao_ms, 3),
            "ao2mo_tflops": round(f_ao2mo / (ao_ms * 1e-3) / 1e12, 3),
            "pipeline_tflops": round((f_ao2mo + f_k) / ((ao_ms + ms_per_step) * 1e-3) / 1e12, 3),
            "rij_ms": round(j_ms, 3),
            "rij_gbs": round(bytes_j / (j_ms * 1e-3) / 1e9, 1),
            "rij_gamma_ms": round(g_ms, 3), "rij_j_ms": round(jj_ms, 3),
            "e_j": e_j,
        })
        log(f"[bench] ao2mo {ao_ms:.2f} ms, RI-J {j_ms:.2f} ms "
            f"({extras['rij_gbs']} GB/s)")

    if not args.no_extras:
        extras["hbm_kernels"] = bench_hbm_kernels(D, dA, nbas, naux, args)
        extras["metric"] = bench_metric(D, naux)

    del dA
    if not args.no_extras and args.batch_mols > 0:
        extras["batched"] = bench_batched(D, O, args)
    if not args.no_extras and args.davidson:
        extras["davidson"] = bench_davidson(D, O, args)

    # ---- CPU baseline: oracle K on a bounded sample of aux slices ----------------------
    cpu = None
    if args.cpu_sample > 0:
        ns = min(args.cpu_sample, naux)
        Bh = np.empty(nocc * nmo * ns)
        from rest_tensors_amd._lib import lib, check
        check(lib().rt_memcpy_d2h(Bh.ctypes.data, dB.ptr, Bh.nbytes))
        Bs = Bh.reshape((nocc, nmo, ns), order="F")
        O.rik_syrk(Bs[:, :, :min(ns, 64)])   # warm the BLAS thread pool
        t0 = time.perf_counter()
        O.rik_syrk(Bs)
        t_cpu = time.perf_counter() - t0
        f_s = float(nmo) * (nmo + 1) * nocc * ns
        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        cpu = {"value": round(f_s / t_cpu / 1e12, 4), "unit": "TFLOP/s", "cores": cores,
               "kind": "port",
               "sample": f"oracle K build (OpenBLAS dsyrk over 64-slice aux chunks, all "
                         f"host threads) over the first {ns} of {naux} aux slices, "
                         f"{t_cpu:.2f} s"}
        log(f"[bench] CPU baseline {cpu['value']} TFLOP/s on {cores} cores ({t_cpu:.2f}s)")

    out = {
        "metric": "RI-K TFLOP/s FP64 vs split-FP32 at ≤1e-8 rel err; "
                  "Davidson sec to 1e-6 resid",
        "metric_detail": "value = FP64 RI-K build K = sum_P B_P^T B_P, credited "
                         "nbas(nbas+1)*nocc*naux flops (upper triangle); reduced-precision "
                         "forms and the Davidson times are under extras",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded ri3ao N(0,1)*exp(-|i-j|/16), orthonormal C from QR)",
        "config": {"workload": f"RI-K build, nbas={nbas} naux={naux} nocc={nocc}",
                   "nbas": nbas, "naux": naux, "nocc": nocc, "seed": args.seed,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3),
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                     "traffic": MEASURED_TRAFFIC_RIK.get((nbas, naux, nocc)),
                     "traffic_unit": "bytes per launch",
                     "algorithmic_bytes": 8.0 * nocc * nmo * naux,
                     "kernel": "dgemm_tn_kernel<128,128,16,2> (rik_mfma)",
                     "kernel_avg_ms": round(k_avg_ms, 4)},
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)


def _timed(D, fn, reps=3):
    """Average event-timed ms of fn() on the default stream (after one warm call)."""
    fn()
    D.synchronize()
    e0, e1 = D.Event(), D.Event()
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    return e0.elapsed_ms(e1) / reps


def bench_hbm_kernels(D, dA, nbas, naux, args):
    """The bandwidth-bound kernels on ri3ao [nbas, nbas, naux] (resident): the
    packed-upper copy [nbas(nbas+1)/2, naux] (rifull_to_matfull_symm), the two
    GEMVs over it that compose J (gamma = A^T d, J = A gamma), the rank-3
    transposes (16 B per element) and the packed-upper unpack of one matrix.
    Bytes are algorithmic (each element read once, written once)."""
    npair = nbas * (nbas + 1) // 2
    out = {}

    def rec(name, ms, nbytes):
        out[name] = {"ms": round(ms, 4), "GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "frac_of_8TBs": round(nbytes / (ms * 1e-3) / (PEAK_HBM_GBS * 1e9), 3)}

    dP = D.DeviceArray(npair * naux)
    ms = _timed(D, lambda: D.rifull_to_matfull_symm(dA, nbas, naux, dP))
    rec("rifull_to_matfull_symm", ms, 16.0 * npair * naux)
    dx = D.DeviceArray(npair)
    D.gen_normal(dx, npair, seed=11)
    dg = D.DeviceArray(naux)
    ms = _timed(D, lambda: D.dgemv("T", npair, naux, 1.0, dP, npair, dx, 0.0, dg))
    rec("dgemv_T_packed", ms, 8.0 * npair * naux)
    ms = _timed(D, lambda: D.dgemv("N", npair, naux, 1.0, dP, npair, dg, 0.0, dx))
    rec("dgemv_N_packed", ms, 8.0 * npair * naux)
    del dP, dx, dg
    dT = D.DeviceArray(nbas * nbas * naux)
    for order in ("jik", "kji"):
        ms = _timed(D, lambda: D.rifull_transpose(dA, nbas, nbas, naux, order, dT), reps=2)
        rec(f"rifull_transpose_{order}", ms, 16.0 * nbas * nbas * naux)
    del dT
    dU = D.DeviceArray(nbas * nbas)
    dUp = D.DeviceArray(npair)
    D.gen_normal(dUp, npair, seed=12)
    ms = _timed(D, lambda: D.upper_unpack(dUp, npair, dU), reps=20)
    rec("upper_unpack", ms, 8.0 * npair + 8.0 * nbas * nbas)
    del dU, dUp
    D.synchronize()
    log("[bench] hbm kernels: " + ", ".join(f"{k} {v['GB/s']} GB/s" for k, v in out.items()))
    return out


def bench_metric(D, naux):
    """The FP64 aux-metric routines on an SPD matrix S = I + G G^T / n (G seeded
    normal) at n = naux: the Cholesky factor (_dpotrf) and the factor of the
    inverse (cholesky_decompose_inverse); at n = naux/2: lapack_power
    (eigen-decomposition, S^-1/2 with eigenvalues below 1e-10 dropped)."""
    from rest_tensors_amd._lib import lib, check
    out = {}
    for n, legs in ((naux, ("dpotrf", "cholesky_decompose_inverse")),
                    (naux // 2, ("lapack_power",))):
        dG, dW = D.DeviceArray(n * n), D.DeviceArray(n * n)
        dS = D.DeviceArray.from_host(np.eye(n).reshape(-1))
        D.gen_normal(dG, n * n, seed=31)
        D.dgemm("N", "T", n, n, n, 1.0 / n, dG, n, dG, n, 1.0, dS, n)
        for leg in legs:
            reps = 1 if leg == "lapack_power" else 3
            D.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                if leg == "dpotrf":
                    check(lib().rt_memcpy_d2d(dW.ptr, dS.ptr, 8 * n * n))
                    assert D.dpotrf("L", n, dW, n) == 0
                elif leg == "cholesky_decompose_inverse":
                    D.cholesky_decompose_inverse("L", n, dS, n, dW, n)
                else:
                    D.matrix_power(n, dS, n, -0.5, 1e-10, dW, n)
            D.synchronize()
            out[leg] = {"n": n, "ms": round((time.perf_counter() - t0) / reps * 1e3, 2)}
        del dG, dS, dW
    log("[bench] metric: " + ", ".join(f"{k} n={v['n']} {v['ms']} ms" for k, v in out.items()))
    return out


def bench_batched(D, O, args):
    """Batched workload: molecules with nbas ~ U{64..384} (seeded), naux = 3 nbas,
    nocc = nbas/8; all stages as grouped launches in one captured HIP graph."""
    from rest_tensors_amd.batch import Batch
    rng = np.random.default_rng(0x5EED0004)
    nbs = rng.integers(64, 385, size=args.batch_mols)
    mols, flops = [], 0.0
    for i, nb in enumerate(nbs):
        nb = int(nb)
        naux, nocc = 3 * nb, max(1, nb // 8)
        dA = D.DeviceArray(nb * nb * naux)
        D.gen_ri3ao(dA, nb, naux, seed=0x5EED0100 + i)
        mols.append(dict(nbas=nb, naux=naux, nocc=nocc, ri3ao=dA,
                         C=D.DeviceArray.from_host(O.gen_orthonormal(nb, seed=0x5EED0200 + i)),
                         D=D.DeviceArray(nb * nb), J=D.DeviceArray(nb * nb),
                         K=D.DeviceArray(nb * nb), B=D.DeviceArray(nocc * nb * naux)))
        flops += 4.0 * nocc * nb * nb * naux + float(nb) * (nb + 1) * nocc * naux
    b = Batch(mols)
    t0 = time.perf_counter()
    b.autotune(reps=1)
    D.synchronize()
    t_tune = time.perf_counter() - t0
    s = D.Stream()
    b.capture(s.handle)
    b.launch(s.handle)
    s.synchronize()
    e0, e1 = D.Event(), D.Event()
    reps = 5
    e0.record(s.handle)
    for _ in range(reps):
        b.launch(s.handle)
    e1.record(s.handle)
    ms = e0.elapsed_ms(e1) / reps
    ej, ex = b.energies(s.handle)
    s.destroy()
    out = {"molecules": len(mols), "nbas_min": int(nbs.min()), "nbas_max": int(nbs.max()),
           "ms_per_batch_graph": round(ms, 3),
           "tflops_ao2mo_plus_k": round(flops / (ms * 1e-3) / 1e12, 3),
           "autotune_s": round(t_tune, 2), "tiles": b.info(),
           "e_x_sum": float(ex.sum()), "e_j_sum": float(ej.sum())}
    log(f"[bench] batched {len(mols)} molecules: {ms:.2f} ms per graph replay "
        f"({out['tflops_ao2mo_plus_k']} TFLOP/s ao2mo+K)")
    del b, mols
    return out


def bench_davidson(D, O, args):
    """Davidson workload: A = diag(eps_a - eps_i) + 2 B B^T, B = 0.05 ri3mo(occ,vir,aux)
    from nbas=512, naux=2048, nocc=64 data (ov = 28672), lowest 8 roots to a 1e-6
    residual (tol 1e-12).  Three preconditioners: the reference's diagonal one
    (plain), the shift-invert correction alone (Woodbury on the low-rank part),
    and the MLP (shift-invert correction in its basis) trained on recorded
    plain-Davidson corrections of two other seeds.  Times include the
    shift-invert setup; training time is reported separately."""
    from oracle import davidson_oracle as DO
    from rest_tensors_amd import davidson as DV
    nbas, naux, nocc, nroots = 512, 2048, 64, 8
    kw = dict(tol=1e-12, maxcyc=1000, maxspace=24)

    def problem(seed):
        dA = D.DeviceArray(nbas * nbas * naux)
        D.gen_ri3ao(dA, nbas, naux, seed=seed)
        dC = D.DeviceArray.from_host(O.gen_orthonormal(nbas, seed=seed + 1))
        dB = D.DeviceArray(nocc * nbas * naux)
        D.ao2mo_occ(dC, nbas, nbas, nocc, dA, naux, dB)
        del dA
        eps = DO.casida_eps(nbas, nocc, seed=seed + 3)
        dvec = (eps[nocc:][None, :] - eps[:nocc][:, None]).reshape(-1, order="F")
        n = dvec.size
        dd = D.DeviceArray.from_host(dvec)
        op = DV.CasidaOp(dB, nocc * nbas, n, naux, 0.05, dd, offset=nocc * nocc)
        hd = D.DeviceArray(n)
        op.hdiag(hd)
        hdh = hd.to_host()
        x0 = np.zeros(n)
        x0[np.argmin(hdh)] = 1.0
        return dict(op=op, hd=hd, x0=D.DeviceArray.from_host(x0), n=n, keep=(dB, dd))

    def solve(p, **extra):
        return DV.solve(p["op"].matvec_fn, p["op"].handle, p["x0"], p["hd"], p["n"], nroots,
                        allow_notconv=True, **kw, **extra)

    def shift_invert(p):
        D.synchronize()
        t0 = time.perf_counter()
        si = DV.ShiftInvert(p["op"], shift=0.05)
        D.synchronize()
        return si, time.perf_counter() - t0

    def leg(res, setup=0.0):
        return {"seconds": round(res.stats["seconds"] + setup, 4),
                "solve_seconds": round(res.stats["seconds"], 4),
                "iterations": res.stats["iterations"], "matvecs": res.stats["matvecs"],
                "converged": all(res.conv), "max_residual": res.stats["max_residual"]}

    target = problem(0x5EED0002)
    # the operator and the shift-invert correction on blocks of k vectors:
    # each pass over B (n x naux, FP64) reads it once
    n_t = target["n"]
    kern = {}
    si_k, _ = shift_invert(target)
    for k in (1, 8):
        dX = D.DeviceArray(n_t * k)
        D.gen_normal(dX, n_t * k, seed=21)
        dY = D.DeviceArray(n_t * k)
        ms = _timed(D, lambda: target["op"].apply(dX, dY, k), reps=10)
        kern[f"casida_matvec_k{k}"] = {"ms": round(ms, 4), "GB/s": round(
            2 * 8.0 * n_t * naux / (ms * 1e-3) / 1e9, 1)}
        ms = _timed(D, lambda: si_k.apply(dX, dY, k), reps=10)
        kern[f"shift_invert_apply_k{k}"] = {"ms": round(ms, 4), "GB/s": round(
            2 * 8.0 * n_t * naux / (ms * 1e-3) / 1e9, 1)}
        del dX, dY
    del si_k
    solve(target)  # warm (kernels, scratch)
    plain = solve(target)
    si_t, t_si = shift_invert(target)
    solve(target, precond="aux", aux=si_t)  # warm
    si_t, t_si = shift_invert(target)
    sionly = solve(target, precond="aux", aux=si_t)
    # training data: plain runs on two other seeds, recorded with their subspaces
    net = DV.Mlp(seed=1)
    t0 = time.perf_counter()
    nsamp = 0
    for seed in (0x5EED1002, 0x5EED2002):
        p = problem(seed)
        ref = solve(p)
        rec = DV.Recorder(p["n"], 4000, subspace_doubles=1 << 29)
        solve(p, recorder=rec)
        ns = rec.count
        nsamp += ns
        tau = D.DeviceArray(p["n"] * ns)
        wts = DV.mlp_targets(rec.x, rec.root[:ns], ns, ref.x, p["n"], tau, recorder=rec)
        si_p, _ = shift_invert(p)
        Sa = D.DeviceArray(p["n"] * ns)
        si_p.apply(rec.r, Sa, ns)
        net.train(rec.r, rec.x, p["hd"], rec.e_k[:ns], rec.e_0[:ns], tau, wts, p["n"], ns,
                  steps=300, batch=32, lr=3e-3, S=Sa, recorder=rec)
        del rec, tau, Sa, si_p
    D.synchronize()
    t_train = time.perf_counter() - t0
    solve(target, precond="mlp", mlp=net, aux=si_t)  # warm
    si_t, t_si = shift_invert(target)
    learned = solve(target, precond="mlp", mlp=net, aux=si_t)
    out = {"n": target["n"], "nroots": nroots, "tol": kw["tol"], "maxspace": kw["maxspace"],
           "plain": leg(plain),
           "shift_invert_only": leg(sionly, t_si),
           "mlp": leg(learned, t_si),
           "shift_invert_setup_s": round(t_si, 4),
           "mlp_over_plain_time": round((learned.stats["seconds"] + t_si) /
                                        plain.stats["seconds"], 3),
           "shift_invert_over_plain_time": round((sionly.stats["seconds"] + t_si) /
                                                 plain.stats["seconds"], 3),
           "training_s": round(t_train, 2), "training_samples": nsamp,
           "kernels": kern,
           "max_eig_diff": float(max(np.abs(plain.e - learned.e).max(),
                                     np.abs(plain.e - sionly.e).max()))}
    log(f"[bench] davidson n={target['n']}: plain {out['plain']['seconds']} s "
        f"({out['plain']['iterations']} it), shift-invert {out['shift_invert_only']['seconds']} s "
        f"({out['shift_invert_only']['iterations']} it), mlp {out['mlp']['seconds']} s "
        f"({out['mlp']['iterations']} it), ratio {out['mlp_over_plain_time']}")
    return out


if __name__ == "__main__":
    main()
