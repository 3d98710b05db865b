"""Benchmark of the CCSD-gradient hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config grad|eri|batch|precision|gemm|int2e] [--waters N]

--config grad (default): the full CCSD gradient of a seeded water octamer
(H2O)8 / cc-pVDZ (192 AOs, nocc 40, nvir 152; rigid monomers, random
orientations, O-O 2.8 +- 0.2 A). RHF and the CCSD T-solve run on the GPU first
and are not timed. One step = lambda solve + dE/dR from the converged
amplitudes with everything resident in HBM (the reference's
solve_lambda_gpu + compute_ml_ccsd_forces_gpu). `value` is the mixed path's
wall time per gradient (split-BF16 MFMA lambda/RDM contractions with FP64
residual refinement, fused int2e_ip1 x Gamma); the plain FP64 path (FP64
lambda, reference-structured unfused gradient) is timed in the same run and
reported in config, with the max |difference| of the two gradients.

--config eri: the fused int2e_ip1 x G backward pass of the ERI autograd
function on (H2O)4 (one step = one pass), value = ERI' quartets/s.

--config precision: the mixed-precision lambda on (H2O)6 (FP32 and split-BF16,
with and without FP64-residual refinement); errors of l1, l2, the RDMs and the
gradient against the FP64 lambda; value = split-BF16 lambda wall-s.

--config batch: 64 seeded C/N/O/H molecules (3-12 heavy atoms, cc-pVDZ) in one
plan; one step = one Schwarz-screened (1e-12) fused int2e_ip1 x Gamma pass over
the whole batch with a separable density Gamma(D) per molecule (one launch per
angular-momentum class for the whole batch, four streams); value = surviving
quartets/s. config.persistent_register_launch times the same pass as ONE
persistent launch of the register-resident per-class bodies over every class of
every molecule; config.persistent_generic_launch_first_molecules the persistent
launch of the LDS kernel's generic body on the first --persistent-molecules
molecules.

--config int2e: the store-mode drop-ins, int2e and int2e_ip1 over all shells of
(H2O)8 (Cartesian s2kl, --waters N for others); value = int2e quartets/s, with the
bytes written per second and the FP64 flop-model roofline of each.

--config gemm: the split-BF16 NT GEMM of the lambda / RDM contractions on the
(H2O)8 sweep shapes (random operands), each beside the same plain bf16 products
through the vendor BLAS; value = algorithmic TFLOP/s (2 batch M N K per launch)
of the l2 (ac|bd) pair product; the three bf16 products the kernel issues per
split pair are reported apart (hardware_bf16_tflops). --gemm-variant picks the
kernel's block shape, --deterministic turns torch's deterministic algorithms on.

Every line carries `roofline` (the dominant kernel, time from HIP events on the
launch stream) and `cpu_baseline`. In the grad line the dominant kernel is the
split-BF16 GEMM (most of the timed step's kernel time), on the lambda sweep's
l2 (ac|bd) pair product (bf16 MFMA roofline, achieved = 2 batch M N K / time); the
fused ERI' pass's FP64 roofline is in config.eri_prime_roofline, with its rate on the
quartets it evaluates (and the reference-equivalent count beside it). The eri and
batch lines use the ERI kernels' FP64 roofline. The grad line's cpu_baseline is
the FP64 oracle timed on this host's cores on samples with exact multiplicities
(cpu_gradient_baseline); the run exits non-zero when the mixed and plain FP64
gradients differ by more than 1e-6 Eh/Bohr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector (and matrix) peak
METRIC = "CCSD-grad wall-s & ERI' quartets/s, mixed vs FP64 plain, @1e-6 Eh/Bohr"


def log(*a):
    print("#", *a, file=sys.stderr, flush=True)


def timed_fused_pass(mol, gp, scale, reps=3):
    """HIP-event time (ms) of one fused int2e_ip1 x Gamma pass on torch's current stream."""
    import torch
    from ccsd_gpu_amd.hip.grad2e import contract_int2e_ip1
    stream = torch.cuda.current_stream()
    contract_int2e_ip1(mol, gp, scale=scale)
    ev = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        contract_int2e_ip1(mol, gp, scale=scale)
        b.record(stream)
        ev.append((a, b))
    torch.cuda.synchronize()
    return float(np.mean([a.elapsed_time(b) for a, b in ev]))


def roofline(mol, kern_ms):
    from ccsd_gpu_amd import flopcount
    flops, quartets, ref_quartets = flopcount.eri_work(mol, 1)
    achieved = flops / (kern_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None}, flops, quartets, ref_quartets


def bench_int2e(args):
    """The store-mode drop-ins: int2e and int2e_ip1 over all shells of (H2O)N (Cartesian
    s2kl, the layout the gradient's AO integrals use), each timed with HIP events on the
    launch stream into a preallocated output."""
    import torch
    from ccsd_gpu_amd import _lib, flopcount, systems
    mol = systems.water_cluster(args.waters or 8, seed=1, cart=True)
    nao = mol.nao_cart()
    npair = nao * (nao + 1) // 2
    lib = _lib.load()
    plan = _lib.get_plan(mol)
    stream = torch.cuda.current_stream()
    res = {}
    for name, fn, ncomp, deriv in (("int2e", lib.ccsd_int2e_s2kl, 1, 0), ("int2e_ip1", lib.ccsd_int2e_ip1_s2kl, 3, 1)):
        out = torch.empty((ncomp, nao, nao, npair), dtype=torch.float64, device="cuda")

        def call():
            _lib.check(fn(plan, _lib.ptr(out), 0, mol.nbas, 1, _lib.current_stream_ptr()), name)
        for _ in range(max(1, args.warmup)):
            call()                          # first call: kernel choices and work lists
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(args.steps):
            call()
        b.record(stream)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / args.steps
        flops, quartets, refq = flopcount.eri_work(mol, deriv)
        gb = out.numel() * 8 / 1e9
        res[name] = {"ms": ms, "quartets": quartets, "quartets_per_s": quartets / (ms * 1e-3),
                     "output_GB": gb, "write_GBps": gb / (ms * 1e-3),
                     "roofline": {"bound": "valu", "achieved": flops / (ms * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                                  "unit": "TFLOP/s", "frac": flops / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                                  "traffic": None}}
        log(f"{name}: {ms:.2f} ms, {quartets / (ms * 1e-3):.3e} quartets/s, {gb:.2f} GB written "
            f"({gb / (ms * 1e-3):.0f} GB/s), {res[name]['roofline']['achieved']:.2f} TF/s FP64 model")
        del out
    if args.copy_calibration:
        copy_calibration()
    head = res["int2e"]
    return {
        "metric": METRIC, "value": head["quartets_per_s"], "unit": "ERI quartets/s (int2e store, s2kl)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"int2e / int2e_ip1 s2kl store over all shells, (H2O){args.waters or 8}/cc-pVDZ "
                               "Cartesian", "results": res},
        "roofline": head["roofline"], "cpu_baseline": None}


def cpu_eri_rate(mol, nshell=3):
    """Oracle (C, OpenMP) int2e_ip1 on the first `nshell` i-shells: flops/s of the CPU path."""
    from ccsd_gpu_amd import flopcount
    from oracle import integrals as oi
    mc = mol.copy()
    mc.cart = True
    t0 = time.perf_counter()
    oi.int2e_s2kl_cart(mc, 0, nshell, 1)
    dt = time.perf_counter() - t0
    f, _, _ = flopcount.eri_work(mol, 1, 0, nshell, prim_cut=False)   # the oracle evaluates every primitive
    return f / dt, dt, oi.num_threads()


def bench_eri(args):
    import torch
    from ccsd_gpu_amd import flopcount, systems
    from ccsd_gpu_amd.autograd import symmetrize8
    from ccsd_gpu_amd.hip.grad2e import contract_int2e_ip1, pack_s2kl, sph_to_cart_4index
    mol = systems.water_cluster(args.waters or 4, seed=1, cart=False)
    n = mol.nao
    gen = torch.Generator(device="cuda").manual_seed(1)
    g = torch.randn((n, n, n, n), dtype=torch.float64, device="cuda", generator=gen)
    g = symmetrize8(g) * np.sqrt(8.0)
    c2s = torch.as_tensor(mol.cart2sph_coeff(), dtype=torch.float64, device="cuda")
    gp = pack_s2kl(sph_to_cart_4index(g, c2s), halve_diag=True)
    del g
    for _ in range(args.warmup):
        contract_int2e_ip1(mol, gp, scale=-8.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        contract_int2e_ip1(mol, gp, scale=-8.0)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    kern_ms = timed_fused_pass(mol, gp, -8.0)
    roof, flops, quartets, refq = roofline(mol, kern_ms)
    # the same pass as one persistent launch of the generic kernel over every class
    from ccsd_gpu_amd.hip.grad2e import set_persistent
    de_c = contract_int2e_ip1(mol, gp, scale=-8.0).clone()
    set_persistent(mol, True)
    try:
        pers_ms = timed_fused_pass(mol, gp, -8.0, reps=2)
        pers_diff = float((contract_int2e_ip1(mol, gp, scale=-8.0) - de_c).abs().max())
    finally:
        set_persistent(mol, False)
    log(f"one persistent launch {pers_ms:.1f} ms vs per-class launches {kern_ms:.1f} ms (max |diff| {pers_diff:.1e})")
    if args.copy_calibration:
        copy_calibration()
    cpu = None
    if not args.no_cpu:
        rate, dt, nth = cpu_eri_rate(mol)
        cpu_s = flopcount.eri_work(mol, 1, prim_cut=False)[0] / rate
        cpu = {"value": refq / cpu_s, "unit": "ERI' quartets/s", "cores": nth, "kind": "port",
               "sample": f"FP64 C oracle int2e_ip1 over i-shells [0,3) of the same molecule ({dt:.1f} s), "
                         "scaled to the full pass by the flop model"}
    return {
        "metric": METRIC, "value": refq / (ms * 1e-3), "unit": "ERI' quartets/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"fused int2e_ip1 x G backward, (H2O){args.waters or 4}/cc-pVDZ, "
                               "seeded 8-fold-symmetric N(0,1) G",
                   "quartets_evaluated": quartets, "quartets_reference_equivalent": refq,
                   "screening": "none",
                   "persistent_single_launch": {"persistent_ms": pers_ms, "per_class_ms": kern_ms,
                                                "max_abs_diff": pers_diff}},
        "roofline": roof, "cpu_baseline": cpu}


def copy_calibration():
    """A plain 256 MiB device copy (256 MiB read + 256 MiB written): the reference that
    scales the FETCH_SIZE / WRITE_SIZE counters of a profiled run to bytes."""
    import torch
    x = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    del x, y


def _sym_density(n, rng):
    a = rng.standard_normal((n, n))
    return (a + a.T) / np.sqrt(n)


def timed_batch_pass(b, dms, reps):
    """HIP-event time (ms) of one batched fused pass on torch's current stream, and the result."""
    import torch
    stream = torch.cuda.current_stream()
    ev = []
    de = None
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        de = b.grad2e_separable(dms)
        e.record(stream)
        ev.append((a, e))
    torch.cuda.synchronize()
    return float(np.mean([a.elapsed_time(e) for a, e in ev])), de


def persistent_vs_per_class(mols, dms, thr):
    """Screened separable fused pass over `mols` as per-class launches (register kernels,
    four streams) and as one persistent launch of the LDS kernel's generic body; times
    (HIP events) and max |difference|."""
    import torch
    from ccsd_gpu_amd.batch import MoleculeBatch
    b = MoleculeBatch(mols)
    b.screen(thr)
    b.grad2e_separable(dms)
    cls_ms, de_c = timed_batch_pass(b, dms, 2)
    de_c = torch.cat(de_c).clone()
    b.set_persistent(True, body="generic")
    b.grad2e_separable(dms)       # first call: generic-kernel choices, work items
    pers_ms, de_p = timed_batch_pass(b, dms, 2)
    q = b.last_quartets()
    diff = float((torch.cat(de_p) - de_c).abs().max())
    b.close()
    log(f"{len(mols)} molecules: one persistent launch (generic body) {pers_ms:.1f} ms vs per-class launches "
        f"{cls_ms:.1f} ms ({q} quartets, max |diff| {diff:.1e})")
    return {"molecules": len(mols), "quartets": q, "persistent_generic_ms": pers_ms, "per_class_ms": cls_ms,
            "persistent_generic_quartets_per_s": q / (pers_ms * 1e-3), "max_abs_diff": diff}


def bench_batch(args):
    import torch
    from ccsd_gpu_amd import flopcount, systems
    from ccsd_gpu_amd.batch import MoleculeBatch
    nmol = args.molecules
    mols = systems.cnoh_batch(nmol, seed=0)
    rng = np.random.default_rng(0)
    dms = [_sym_density(m.nao_nr(), rng) for m in mols]
    b = MoleculeBatch(mols)
    log(f"{nmol} C/N/O/H molecules, nao {min(m.nao_nr() for m in mols)}-{max(m.nao_nr() for m in mols)}, "
        f"{sum(m.natm for m in mols)} atoms, {int(b.mol_nbas.sum())} shells")
    b.screen(args.thr)
    t0 = time.perf_counter()
    for _ in range(args.warmup):
        b.grad2e_separable(dms)       # first call: Schwarz factors, work lists, per-class kernel choice
    torch.cuda.synchronize()
    log(f"warmup (incl. screening set-up and autotuning) {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        de = b.grad2e_separable(dms)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    de_s = torch.cat(de).clone()
    surv = b.last_quartets()
    kern_ms, _ = timed_batch_pass(b, dms, 2)
    # the same pass as ONE persistent launch of the register-resident bodies over every
    # class of every molecule (plus one generic-body launch for the classes without a
    # register body), against the per-class launch set above
    b.set_persistent(True, body="register", order=args.persistent_order)
    b.grad2e_separable(dms)           # first call: work items, each item's duration recorded
    reg_ms, de_r = timed_batch_pass(b, dms, 1)   # items cut again from those durations
    launches = b.last_launches()
    reg_diff = float((torch.cat(de_r) - de_s).abs().max())
    b.set_persistent(False)
    n_reg = sum(1 for r in launches if r["kernel"] == 3)
    log(f"one persistent launch (register bodies, {args.persistent_order} order, {n_reg} of {len(launches)} "
        f"classes; the rest in one generic-body launch) {reg_ms:.1f} ms vs per-class launches {kern_ms:.1f} ms "
        f"(ratio {reg_ms / kern_ms:.2f}, max |diff| {reg_diff:.1e})")
    q = b.schwarz().cpu().numpy()
    flops = 0.0
    s0 = 0
    for m in mols:
        f, _ = flopcount.eri_work_screened(m, 1, q[s0:s0 + m.nbas, s0:s0 + m.nbas], args.thr)
        flops += f
        s0 += m.nbas
    achieved = flops / (kern_ms * 1e-3) / 1e12
    roof = {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None}
    # the persistent launch with the LDS kernel's generic wave-per-quartet body against the
    # per-class launches, on the first few molecules (it would take minutes on all)
    pers = persistent_vs_per_class(mols[:args.persistent_molecules], dms[:args.persistent_molecules], args.thr)
    # the same pass unscreened, for the error and gain of screening
    b.screen(0.0)
    b.grad2e_separable(dms)
    full_ms, de_f = timed_batch_pass(b, dms, 1)
    err = float((de_s - torch.cat(de_f)).abs().max())
    net = max(float(d.sum(0).abs().max()) for d in de_f)
    refq = 0
    for m in mols:
        ndc = int(m._bas[:, 3].sum())
        refq += ndc * ndc * ndc * (ndc + 1) // 2
    cpu = None
    if not args.no_cpu:
        rate, dt, nth = cpu_eri_rate(mols[-1])
        cpu_s = sum(flopcount.eri_work(m, 1, prim_cut=False)[0] for m in mols) / rate   # no screening, no cut
        cpu = {"value": b.quartets_unscreened() / cpu_s, "unit": "ERI' quartets/s", "cores": nth, "kind": "port",
               "sample": f"FP64 C oracle int2e_ip1 over i-shells [0,3) of one molecule of the batch ({dt:.1f} s, "
                         f"{rate / 1e9:.1f} GFLOP/s), scaled by the flop model to the whole batch, "
                         "every quartet and primitive (the CPU path screens nothing)"}
    return {
        "metric": METRIC, "value": surv / (ms * 1e-3), "unit": "ERI' quartets/s (Schwarz survivors)", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched fused int2e_ip1 x Gamma(D) over {nmol} seeded C/N/O/H molecules "
                               "(3-12 heavy atoms, cc-pVDZ), one launch per angular-momentum class for the batch",
                   "schwarz_threshold": args.thr, "quartets_evaluated": surv,
                   "quartets_unscreened": b.quartets_unscreened(), "quartets_reference_equivalent": refq,
                   "unscreened_pass_ms": full_ms, "screened_pass_ms": kern_ms,
                   "persistent_register_launch": {"ms": reg_ms, "per_class_launches_ms": kern_ms,
                                                  "ratio_to_per_class": reg_ms / kern_ms,
                                                  "register_body_classes": n_reg, "classes": len(launches),
                                                  "item_order": args.persistent_order,
                                                  "max_abs_diff_vs_per_class": reg_diff},
                   "persistent_generic_launch_first_molecules": pers,
                   "max_abs_diff_screened_vs_unscreened": err, "max_net_force_unscreened": net},
        "roofline": roof, "cpu_baseline": cpu}


def bench_precision(args):
    """Accuracy and speed of the mixed-precision lambda solve on (H2O)6 / cc-pVDZ
    (nocc 30, nvir 114): l1, l2, the 1- and 2-RDMs and the gradient against the
    FP64 lambda, for FP32 and split-BF16 MFMA contractions, with and without the
    FP64-residual refinement."""
    import torch
    from ccsd_gpu_amd import pipeline, systems
    from ccsd_gpu_amd.gpu_ccsd_grad import _gamma1_intermediates_gpu, _gamma2_intermediates_gpu, grad_elec_gpu
    from ccsd_gpu_amd.gpu_ccsd_lambda import solve_lambda_gpu
    nw = args.waters or 6
    mol = systems.water_cluster(nw, seed=1, cart=False)
    mf, mycc = pipeline.setup(mol, log=log)
    t1, t2 = mycc.t1, mycc.t2

    def solve(prec, refine=True):
        mycc.conv_tol_normt = 1e-8
        mycc.lambda_refine = refine
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        l1, l2 = solve_lambda_gpu(mycc, t1, t2, precision=prec, return_tensors=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        mycc.lambda_refine = True
        return l1.clone(), l2.clone(), dt, mycc._lambda_iterations, mycc._lambda_fp64_updates

    def rdms(l1, l2, prec):
        d1 = _gamma1_intermediates_gpu(mycc, t1, t2, l1, l2, prec)
        d2 = _gamma2_intermediates_gpu(mycc, t1, t2, l1, l2, prec)
        return [x for x in list(d1) + list(d2) if x is not None]

    solve("fp64")                       # warm-up (static caches, library init)
    l1r, l2r, t_ref, it_ref, _ = solve("fp64")
    d_ref = rdms(l1r, l2r, "fp64")
    g_ref = grad_elec_gpu(mycc, t1, t2, l1r, l2r, precision="fp64")
    rows = {"fp64": {"lambda_s": t_ref, "iterations": it_ref}}
    for prec in ("fp32", "bf16x3"):
        for refine in (False, True):
            solve(prec, refine)         # warm-up of this variant
            l1, l2, dt, it, nf = solve(prec, refine)
            d = rdms(l1, l2, prec)
            g = grad_elec_gpu(mycc, t1, t2, l1, l2, precision=prec)
            rows[prec + ("" if refine else "_no_refinement")] = {
                "lambda_s": dt, "iterations": it, "fp64_updates": nf,
                "max_err_l1": float((l1 - l1r).abs().max()), "max_err_l2": float((l2 - l2r).abs().max()),
                "max_err_rdm1": max(float((a - b).abs().max()) for a, b in zip(d[:4], d_ref[:4])),
                "max_err_rdm2": max(float((a - b).abs().max()) for a, b in zip(d[4:], d_ref[4:])),
                "max_err_grad_Eh_per_Bohr": float(np.abs(g - g_ref).max())}
            log(f"{prec} refine={refine}: {json.dumps(rows[prec + ('' if refine else '_no_refinement')])}")
    best = rows["bf16x3"]
    return {
        "metric": METRIC, "value": best["lambda_s"], "unit": "s per lambda solve (split-BF16, refined to 1e-8)",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": best["lambda_s"] * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+bf16x3", "data": "synthetic",
        "config": {"workload": f"CCSD lambda, (H2O){nw}/cc-pVDZ seeded, nocc {mycc.nocc}, nvir {mol.nao - mycc.nocc}, "
                               "errors against the FP64 lambda",
                   "speedup_bf16x3_vs_fp64": t_ref / best["lambda_s"], "results": rows},
        "roofline": None, "cpu_baseline": None}


def _timed(fn):
    t0 = time.perf_counter()
    out = fn()
    return time.perf_counter() - t0, out


def cpu_gradient_baseline(mol, nocc, lambda_iters, zvector_iters):
    """The FP64 CPU restatement of the timed gradient path (oracle/), timed on this
    host's cores, on bounded samples whose multiplicities are exact:

      * int2e_ip1 and int2e (the reference's per-block ERI' and ERI passes) over the
        i-shells of the first water: every water of the cluster carries the same
        shells and the oracle screens nothing, so the full pass is 8 times the sample
        (n_waters times in general);
      * the lambda intermediates once and one lambda update (numpy, the oracle's
        equations at the real o / v sizes, operands random) times the plain GPU
        path's iteration count;
      * the 1- and 2-RDM intermediates and the MO -> AO transform of the 2-RDM once;
      * one J/K build (the Z-vector's fvind and the two vhf builds) times the GPU
        solve's CG iterations + 2;
      * the one-electron derivative integrals once.
    A one-thread figure comes from the same ERI' sample and a lambda-sized GEMM timed
    with one thread, applied to the all-thread total as measured slowdown factors."""
    from threadpoolctl import threadpool_limits
    from oracle import chem, chem_large
    from oracle import integrals as oi
    nthreads = oi.num_threads()
    mc = mol.copy()
    mc.cart = True
    nw = mol.natm // 3
    b1 = int(np.searchsorted(mol._bas[:, 0], 3))                  # shells of the first water
    parts = {}
    t_ip1, _ = _timed(lambda: oi.int2e_s2kl_cart(mc, 0, b1, 1))
    t_eri, _ = _timed(lambda: oi.int2e_s2kl_cart(mc, 0, b1, 0))
    parts["int2e_ip1"] = t_ip1 * nw
    parts["int2e"] = t_eri * nw
    log(f"cpu baseline: ip1 sample {t_ip1:.1f} s, int2e sample {t_eri:.1f} s (x{nw})")
    n = mol.nao_nr()
    v = n - nocc
    rng = np.random.default_rng(7)

    class _E:
        pass

    eris = _E()
    eris.fock = np.diag(np.concatenate([-1.0 - rng.random(nocc), 0.5 + rng.random(v)]))
    eris.mo_energy = eris.fock.diagonal().copy()
    for name, shape in (("oooo", (nocc,) * 4), ("ovoo", (nocc, v, nocc, nocc)), ("ovvo", (nocc, v, v, nocc)),
                        ("oovv", (nocc, nocc, v, v)), ("ovov", (nocc, v, nocc, v)), ("ovvv", (nocc, v, v, v)),
                        ("vvvv", (v, v, v, v))):
        # (the v^4 block is filled, not drawn: drawing 4 GB of normals costs more than the update)
        setattr(eris, name, np.full(shape, 0.01) if name == "vvvv" else rng.standard_normal(shape) * 0.01)
    t1 = rng.standard_normal((nocc, v)) * 0.01
    t2 = rng.standard_normal((nocc, nocc, v, v)) * 0.01
    t_imds, imds = _timed(lambda: chem.make_intermediates(t1, t2, eris))
    t_upd, _ = _timed(lambda: chem.update_lambda(t1, t2, t1, t2, eris, imds))
    parts["lambda"] = t_imds + t_upd * lambda_iters
    del imds
    log(f"cpu baseline: lambda intermediates {t_imds:.1f} s, one update {t_upd:.1f} s (x{lambda_iters})")
    t_r1, _ = _timed(lambda: chem.gamma1_intermediates(t1, t2, t1, t2))
    t_r2, d2 = _timed(lambda: chem.gamma2_intermediates(t1, t2, t1, t2))
    del eris
    c = np.linalg.qr(rng.standard_normal((n, n)))[0]
    t_ao, _ = _timed(lambda: chem_large._symmetrized_ao_rdm2([d2], c))
    parts["rdm"] = t_r1 + t_r2
    parts["mo2ao"] = t_ao
    log(f"cpu baseline: RDMs {t_r1 + t_r2:.1f} s, MO->AO {t_ao:.1f} s")
    mf = chem_large.RHF(mol)
    mf._eri = np.full((n, n, n, n), 0.01)
    dm = rng.standard_normal((n, n))
    dm = dm + dm.T
    t_jk, _ = _timed(lambda: mf.get_jk(dm))
    del mf
    parts["jk"] = t_jk * (zvector_iters + 2)
    t_1e = sum(_timed(lambda k=k: oi.int1e(mol, k))[0] for k in ("ipovlp", "ipkin", "ipnuc_atom"))
    parts["one_electron"] = t_1e
    total = sum(parts.values())
    # one thread: the same ERI' sample, and a GEMM of the lambda update's shape class
    oi.set_num_threads(1)
    try:
        t_ip1_1, _ = _timed(lambda: oi.int2e_s2kl_cart(mc, 0, 1, 1))
    finally:
        oi.set_num_threads(nthreads)
    t_ip1_n, _ = _timed(lambda: oi.int2e_s2kl_cart(mc, 0, 1, 1))
    a = rng.standard_normal((nocc * nocc // 2, v * v))
    bmat = rng.standard_normal((v * v, v * v // 16))
    t_g_n, _ = _timed(lambda: a @ bmat)
    with threadpool_limits(limits=1):
        t_g_1, _ = _timed(lambda: a @ bmat)
    eri_share = (parts["int2e_ip1"] + parts["int2e"]) / total
    slow = eri_share * (t_ip1_1 / t_ip1_n) + (1 - eri_share) * (t_g_1 / t_g_n)
    return {"value": total, "unit": "s per CCSD gradient", "cores": nthreads, "kind": "port",
            "sample": (f"FP64 oracle timed on {nthreads} host threads: C int2e_ip1 and int2e over the first "
                       f"water's i-shells x {nw} (exact: identical shells, nothing screened), numpy lambda "
                       f"intermediates + one update x {lambda_iters} iterations, RDMs and MO->AO once, one J/K "
                       f"x {zvector_iters + 2}, 1e derivatives once; random operands of the real sizes"),
            "parts_s": {k: round(x, 3) for k, x in parts.items()},
            "one_thread": {"value": total * slow, "unit": "s per CCSD gradient", "cores": 1,
                           "sample": (f"the all-thread total scaled by measured one-thread slowdowns: ERI' "
                                      f"{t_ip1_1 / t_ip1_n:.1f}x (first i-shell), GEMM {t_g_1 / t_g_n:.1f}x "
                                      f"({a.shape[0]}x{a.shape[1]}x{bmat.shape[1]})")}}


def bench_grad(args):
    import torch
    from ccsd_gpu_amd import pipeline, systems
    nw = args.waters or 8
    mol = systems.water_cluster(nw, seed=1, cart=False)
    log(f"(H2O){nw}/cc-pVDZ nao={mol.nao} nbas={mol.nbas}")
    t0 = time.perf_counter()
    mf, mycc = pipeline.setup(mol, log=log)
    torch.cuda.synchronize()
    log(f"setup (RHF {mf.e_tot:.10f}, CCSD corr {mycc.e_corr:.10f}) {time.perf_counter() - t0:.1f} s")
    res = []
    first = None      # the first gradient on this plan and CCSD object (cold: kernel choices timed,
    #                   lambda update graph captured, vvvv pair matrices built), when warm-up ran
    # two small fills (12345 int16 elements, a kernel nothing else launches) bracket each
    # timed step outside its timed region, so a kernel trace of this run can cut out one
    # mixed gradient
    marker = torch.empty(12345, dtype=torch.int16, device="cuda")
    for s in range(args.warmup + args.steps):
        if s >= args.warmup:
            marker.fill_(1)
        de, info = pipeline.ccsd_gradient(mf, mycc, "mixed", args.precision)
        if s >= args.warmup:
            marker.fill_(2)
        log(f"mixed step {s}: {json.dumps(info)}")
        hist = getattr(mycc, "_lambda_history", [])
        log("lambda history: " + " ".join(f"{p[:2]}{'c' if c else ''}:{n:.1e}" for p, c, n in hist))
        if s >= args.warmup:
            res.append((de, info))
        elif s == 0:
            first = info
    wall = float(np.mean([i["total_s"] for _, i in res]))
    de_mixed = res[-1][0]
    # the plain path warmed once too (its first call builds the per-block work lists)
    for rep in range(2):
        de_plain, info_plain = pipeline.ccsd_gradient(mf, mycc, "plain")
        log(f"plain {'warm-up' if rep == 0 else 'timed'}: {json.dumps(info_plain)}")
    diff = float(np.abs(de_mixed - de_plain).max())
    net_force = float(np.abs(de_mixed.sum(axis=0)).max())
    coords = mol.atom_coords()
    torque = float(np.abs(np.cross(coords, de_mixed).sum(axis=0)).max())

    # dominant ERI kernel: the fused ip1 x Gamma pass, timed live with HIP events
    nc = mol.nao_cart()
    gen = torch.Generator(device="cuda").manual_seed(3)
    gp = torch.randn((nc, nc, nc * (nc + 1) // 2), dtype=torch.float64, device="cuda", generator=gen)
    kern_ms = timed_fused_pass(mol, gp, -2.0, reps=2)
    del gp
    roof, flops, quartets, refq = roofline(mol, kern_ms)
    # the lambda sweeps' largest product (l2 (ac|bd), i <= j) on the MFMA roofline
    _, gm, gn, gk, gb = LAMBDA_GEMMS[0]
    gemm_ms = timed_gemm(gm, gn, gk, reps=3, nb=gb)
    setup_s = getattr(mf, "_setup_times", {})

    cpu = None
    if not args.no_cpu:
        cpu = cpu_gradient_baseline(mol, mycc.nocc, info_plain["lambda_iterations"] or 20,
                                    info_plain.get("zvector_iterations") or 30)
        log(f"cpu baseline: {json.dumps(cpu)}")

    gate_ok = diff <= 1e-6 and net_force <= 1e-7
    line = {
        "metric": METRIC, "value": wall, "unit": "s per CCSD gradient (mixed path)", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+bf16x3", "data": "synthetic",
        "config": {"workload": f"CCSD gradient (lambda + dE/dR), (H2O){nw}/cc-pVDZ seeded, nocc {mycc.nocc}, "
                               f"nvir {mol.nao - mycc.nocc}",
                   "mixed_precision": args.precision,
                   "mixed": res[-1][1], "plain_fp64": info_plain,
                   "plain_fp64_s": info_plain["total_s"],
                   "speedup_mixed_vs_plain": info_plain["total_s"] / wall,
                   "max_abs_diff_mixed_vs_plain_Eh_per_Bohr": diff,
                   "net_force": net_force, "net_torque": torque, "accuracy_gate_1e-6_ok": gate_ok,
                   "first_gradient_cold": None if first is None else {
                       "total_s": first["total_s"], "lambda_phases_s": first["lambda_phases_s"],
                       "phases_s": {k: round(v, 4) for k, v in first["phases_s"].items()}},
                   "outside_timed_region_s": {k: (round(v, 3) if isinstance(v, float) else v)
                                              for k, v in setup_s.items()},
                   "ao_eris_note": "the Cartesian AO ERIs are computed once (int2e, register store kernels) "
                                   "before the timed region and reused by the SCF, the MO integrals and the "
                                   "gradient's Imat / J / K; the reference's gradient evaluates int2e per "
                                   "block inside its loop",
                   "eri_prime_quartets_per_s": quartets / (kern_ms * 1e-3),
                   "eri_prime_quartets_evaluated": quartets,
                   "eri_prime_reference_equivalent_quartets_per_s": refq / (kern_ms * 1e-3),
                   "quartets_reference_equivalent": refq,
                   "eri_prime_pass_ms": kern_ms,
                   "eri_prime_roofline": roof,
                   "lambda_gemm": {"shape_mnk": [gm, gn, gk], "batch": gb, "ms": gemm_ms,
                                   "hardware_bf16_tflops": 3 * 2.0 * gb * gm * gn * gk / (gemm_ms * 1e-3) / 1e12}},
        "roofline": gemm_roofline(gm, gn, gk, gemm_ms, gb), "cpu_baseline": cpu}
    return line


BF16_PEAK_TFLOPS = 2500.0   # MI355X dense bf16 MFMA peak

# the split-BF16 products of one (H2O)8 lambda sweep (nocc 40, nvir 152): (name, M, N, K, batch).
# The l2 (ac|bd) term runs as its symmetric and antisymmetric pair halves, one batch of two
# products over the v(v+1)/2 = 11628 index pairs (K padded to 11632 for 16-byte rows).
LAMBDA_GEMMS = [("l2 (ac|bd) pair halves, i<=j", 820, 11628, 11632, 2), ("ring terms (ov x ov)", 6080, 6080, 6080, 1),
                ("l2t1 ovvv", 1600, 23104, 6080, 1), ("l2tau voov / woooo l2", 1600, 23104, 1600, 1)]


def timed_gemm(m, n, k, reps=5, seed=0, nb=1):
    """HIP-event time (ms) of one split-BF16 NT GEMM launch (batch nb, no K split) on random planes."""
    import torch
    from ccsd_gpu_amd import mfma
    gen = torch.Generator(device="cuda").manual_seed(seed)
    A = mfma.split(torch.randn((nb, m, k), dtype=torch.float64, device="cuda", generator=gen))
    B = mfma.split(torch.randn((nb, n, k), dtype=torch.float64, device="cuda", generator=gen))
    out = torch.empty((nb, m, n), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    mfma.gemm_nt(A, B, out=out, ksplit=1)
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        mfma.gemm_nt(A, B, out=out, ksplit=1)
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def timed_vendor_bf16(m, n, k, reps=5, seed=0):
    """HIP-event time (ms) of ONE plain bf16 product of the same shape through torch.mm
    (the vendor BLAS): the scale against which the three-product split kernel is read."""
    import torch
    gen = torch.Generator(device="cuda").manual_seed(seed)
    a = torch.randn((m, k), dtype=torch.float32, device="cuda", generator=gen).to(torch.bfloat16)
    b = torch.randn((k, n), dtype=torch.float32, device="cuda", generator=gen).to(torch.bfloat16).contiguous()
    c = torch.mm(a, b)
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        torch.mm(a, b, out=c)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


GEMM_TRAFFIC = os.path.join(ROOT, "profiles", "gemm_bf16x3_traffic.json")


def gemm_roofline(m, n, k, ms, nb=1):
    """MFMA roofline of a split-BF16 GEMM launch on its algorithmic work, 2 nb M N K
    (the three bf16 products the kernel issues per element pair are reported apart,
    as hardware_bf16_tflops). `traffic` is the HBM bytes per launch of this shape
    measured by rocprofv3 FETCH_SIZE + WRITE_SIZE (copy-calibrated) in the committed
    profile, when it has a row for the shape."""
    achieved = 2.0 * nb * m * n * k / (ms * 1e-3) / 1e12
    traffic = None
    if os.path.exists(GEMM_TRAFFIC):
        for row in json.load(open(GEMM_TRAFFIC)).get("shapes", []):
            if [row.get("m"), row.get("n"), row.get("k"), row.get("batch", 1)] == [m, n, k, nb]:
                traffic = row.get("hbm_bytes")
    return {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / BF16_PEAK_TFLOPS, "traffic": traffic,
            "hardware_bf16_tflops": 3 * achieved}


def bench_gemm(args):
    """The split-BF16 GEMM of the lambda / RDM contractions on the (H2O)8 sweep shapes."""
    from ccsd_gpu_amd import _lib
    _lib.check(_lib.load().ccsd_gemm_set_variant(args.gemm_variant), "ccsd_gemm_set_variant")
    rows = []
    for name, m, n, k, nb in LAMBDA_GEMMS:
        ms = timed_gemm(m, n, k, reps=max(1, args.steps), nb=nb)
        roof = gemm_roofline(m, n, k, ms, nb)
        vms = nb * timed_vendor_bf16(m, n, k, reps=max(1, args.steps))
        rows.append({"product": name, "m": m, "n": n, "k": k, "batch": nb, "ms": ms,
                     "algorithmic_tflops": roof["achieved"], "frac": roof["frac"],
                     "hardware_bf16_tflops": roof["hardware_bf16_tflops"],
                     "vendor_bf16_mm_ms": vms, "vendor_bf16_tflops": 2.0 * nb * m * n * k / vms / 1e9})
        log(f"{name}: {nb} x {m}x{n}x{k} {ms:.3f} ms, {roof['achieved']:.0f} TF/s algorithmic ({roof['frac']:.1%} "
            f"of the bf16 peak), {roof['hardware_bf16_tflops']:.0f} TF/s of bf16 MFMA work; the same plain bf16 "
            f"products through the vendor library {vms:.3f} ms ({2.0 * nb * m * n * k / vms / 1e9:.0f} TF/s)")
    head = rows[0]
    return {
        "metric": METRIC, "value": head["algorithmic_tflops"], "unit": "TFLOP/s (2 batch M N K per launch)",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": head["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3 (fp32 accumulate, fp64 out)", "data": "synthetic",
        "config": {"workload": "split-BF16 NT GEMM on the (H2O)8/cc-pVDZ lambda-sweep shapes, random operands",
                   "gemm_variant": args.gemm_variant, "products": rows},
        "roofline": gemm_roofline(head["m"], head["n"], head["k"], head["ms"], head["batch"]), "cpu_baseline": None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="grad", choices=["grad", "eri", "batch", "precision", "gemm", "int2e"])
    ap.add_argument("--molecules", type=int, default=64, help="batch config: number of molecules")
    ap.add_argument("--thr", type=float, default=1e-12, help="batch config: Schwarz threshold")
    ap.add_argument("--persistent-molecules", type=int, default=4,
                    help="batch config: molecules of the generic-body persistent-launch comparison")
    ap.add_argument("--persistent-order", default="class", choices=["class", "cost"],
                    help="batch config: work-item order of the register-body persistent launch")
    ap.add_argument("--waters", type=int, default=0)
    ap.add_argument("--precision", default="bf16x3", choices=["bf16x3", "fp32"])
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with a plain 256 MiB device copy (HBM counter calibration)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--deterministic", action="store_true",
                    help="torch's deterministic algorithms on (bitwise-repeatable gradients)")
    ap.add_argument("--gemm-variant", type=int, default=0,
                    help="gemm config: block shape of the split-BF16 GEMM (ccsd_gemm_set_variant)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from ccsd_gpu_amd import _lib
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    _lib.load()
    from ccsd_gpu_amd.determinism import deterministic
    with deterministic(args.deterministic):
        line = {"grad": bench_grad, "eri": bench_eri, "batch": bench_batch, "precision": bench_precision,
                "gemm": bench_gemm, "int2e": bench_int2e}[args.config](args)
    if args.deterministic:
        line["config"]["deterministic"] = True
    if args.copy_calibration and args.config != "eri":
        copy_calibration()
    print(json.dumps(line), flush=True)
    if line.get("config", {}).get("accuracy_gate_1e-6_ok") is False:
        print("bench.py: mixed and plain FP64 gradients differ by more than 1e-6 Eh/Bohr or the net force is "
              "not zero; the timing is invalid", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
