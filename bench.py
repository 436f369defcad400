"""Benchmark of the NQS VMC hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c5]
                    [--budget K] [--precision mixed|fp64|fp32] [--gram direct|factored|auto]
                    [--move flip|exchange|learned] [--no-cpu] [--c3]

A step is one full SR iteration (Metropolis sampling of every chain, local
energies, O_k, MinSR / CG-SR solve, parameter update) through the library's
C ABI.  Default workload: the metric's configuration, kagome Heisenberg on
the 2x4-cell periodic cluster (N = 24, Sz = 0), complex RBM alpha = 4
(M = 96), 2048 chains x 8 = 16384 samples per iteration, MinSR, run twice:
with plain exchange Metropolis and with the learned autoregressive proposal
(Adam lr 1e-3, retrained every 10 iterations on the last 16384 samples).
Each leg first optimises its own wavefunction for the workload's budget of
iterations (C5: 300 holomorphic SR iterations, untimed) and reports the
energy error against exact diagonalisation (E0 = -10.759897248084), then
times K MinSR iterations at that state.  With --c3 the line also carries the
J1-J2 10x10 SR-iteration time against the reference CPU (configuration 3);
`--workload c3` benches that configuration on its own.  Prints one JSON
line.

value = effective samples per second of the sampling phase (the learned
leg for C5): sum over chains of n_c / (2 tau_int) with tau_int from the
pooled per-chain autocorrelation of E_loc (Sokal window, c = 5; 95%
interval from a jackknife over chain groups), divided by the sampling-stage
time of the timed iterations (HIP events on the library's stream).
ess_per_s_iteration divides by the whole SR-iteration wall time instead;
ms_per_step is the SR-iteration wall time.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from spin_based_neural_network_amd import abi  # noqa: E402
from spin_based_neural_network_amd.stats import pooled_tau, tau_ci  # noqa: E402

METRIC = "effective samples/s & SR-iter wall time, learned vs plain MH, same ΔE vs ED"
PEAK_FP64 = 78.6e12          # MI355X dense FP64 (vector / MFMA), FLOP/s
PEAK_BF16 = 2.5e15           # MI355X dense BF16 MFMA, FLOP/s
PEAK_HBM = 8.0e12            # MI355X HBM3E, B/s
LEARNED_LR = 1e-3            # Adam rate of the learned proposal (configuration 5)

WORKLOADS = {
    # 6x6 periodic Heisenberg + Marshall, complex RBM alpha=4, MinSR
    "c2": dict(desc="Heisenberg AFM 6x6 PBC, Marshall sign, Sz=0; complex RBM alpha=4 (M=144, "
                    "P=10728 real); 1024 chains x 8 samples = 8192/iter; exchange moves, "
                    "1 sweep between samples; MinSR shift 1e-4",
               ham=1, Lx=6, Ly=6, kind=6, M=144, chains=1024, per_chain=8, move=1, sz0=1,
               periodic=1, marshall=True, method="minsr", shift=1e-4, lr=1e-2, ed=None),
    # 1D TFIM N=16 PBC, RBM alpha=2, CG-SR
    "c1": dict(desc="TFIM chain N=16 PBC J=h=1; RBM alpha=2 (M=32, P=560); 64 chains x 16 = "
                    "1024/iter; single flips; CG-SR shift 1e-3 lr 0.01",
               ham=0, Lx=16, Ly=1, kind=5, M=32, chains=64, per_chain=16, move=0, sz0=0,
               periodic=1, marshall=False, method="cg", shift=1e-3, lr=1e-2,
               ed=-2.0 / math.sin(math.pi / 32)),
    # J1-J2 10x10 PBC J2=0.5, RBM alpha=8, MinSR 16384^2 Gram
    "c3": dict(desc="J1-J2 10x10 PBC J2/J1=0.5, Sz=0; RBM alpha=8 (M=800, P=80900); 4096 "
                    "chains x 4 = 16384/iter; exchange moves; MinSR 16384^2 Gram",
               ham=2, Lx=10, Ly=10, kind=5, M=800, chains=4096, per_chain=4, move=1, sz0=1,
               periodic=1, marshall=False, method="minsr", shift=1e-4, lr=1e-2, ed=None),
    # kagome 2x4 PBC (N=24), complex RBM M=96, MinSR
    "c5": dict(desc="kagome Heisenberg 2x4 cells PBC (N=24), J=1, Sz=0; complex RBM M=96 "
                    "(P=4848 real); 2048 chains x 8 = 16384/iter; MinSR shift 1e-4 lr 0.01; "
                    "learned AR-MLP proposal 24-64-64-24 (BF16, Adam lr 1e-3, retrained every "
                    "10 iterations on the last 16384 samples, MH-corrected) against plain "
                    "exchange Metropolis (1 sweep between samples)",
               ham=6, Lx=2, Ly=4, kind=6, M=96, chains=2048, per_chain=8, move=1, sz0=1,
               periodic=1, marshall=False, method="minsr", shift=1e-4, lr=1e-2,
               ed=-10.759897248084, opt_method="holo", opt_lr=2e-2, opt_shift=1e-3, budget=300),
    # the same system optimised as the reference's kagome runs do (holomorphic
    # SR, lr 0.02, shift 1e-3: scripts/research_kagome_N12_convergence.c:68-69);
    # the real-projected MinSR of c5 cannot move the phases of psi
    "c5h": dict(desc="kagome Heisenberg 2x4 cells PBC (N=24), Sz=0; complex RBM M=96; 2048 "
                     "chains x 8 = 16384/iter; exchange moves; holomorphic CG-SR lr 0.02 shift 1e-3",
                ham=6, Lx=2, Ly=4, kind=6, M=96, chains=2048, per_chain=8, move=1, sz0=1,
                periodic=1, marshall=False, method="holo", shift=1e-3, lr=2e-2,
                ed=-10.759897248084),
    # 32 independent TFIM chains of uneven size, one batch (nqs_hip_multi_*)
    "c4": dict(desc="32 TFIM chains PBC, J=1, N_i in {16,24,32,48,64,96,128} and h_i ~ U[0.5,1.5] "
                    "(seed 0xC4); RBM alpha=2 each; 256 chains x 8 = 2048 samples per instance; "
                    "single flips, 1 sweep between samples; CG-SR shift 1e-3 lr 0.01 (100 / 1e-6); "
                    "all instances in one launch sequence, one hipGraph per iteration",
               multi=True, n=32, sizes=(16, 24, 32, 48, 64, 96, 128), chains=256, per_chain=8,
               shift=1e-3, lr=1e-2, seed=0xC4, method="cg", kind=5, ham=0),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(lib, w, args, move=None):
    """Ansatz (random init N(0, 0.01)) and sampler of the workload; `move`
    overrides the workload's move kind ("flip", "exchange", "learned")."""
    N = 3 * w["Lx"] * w["Ly"] if w["ham"] == 6 else w["Lx"] * w["Ly"]
    cfg = abi.config_defaults(ansatz=w["kind"], hamiltonian=w["ham"], rbm_hidden_units=w["M"],
                              j_coupling=1.0, j2_coupling=0.5, transverse_field=1.0, kagome_pbc=1,
                              num_samples=w["chains"] * w["per_chain"], num_thermalize=4 * N,
                              num_decorrelate=N, rng_seed=20260,
                              learning_rate=args.lr if getattr(args, "lr", None) else w["lr"],
                              sr_diag_shift=w["shift"], sr_cg_max_iters=100, sr_cg_tol=1e-6)
    a = lib.nqs_ansatz_create(ctypes.byref(cfg), N)
    if not a:
        raise RuntimeError("nqs_ansatz_create failed")
    P = lib.nqs_ansatz_num_params(a)
    par = np.ctypeslib.as_array(ctypes.cast(lib.nqs_ansatz_params_raw(a), ctypes.POINTER(ctypes.c_double)), (P,))
    par[:] = np.random.default_rng(12345).normal(0.0, 0.01, P)
    fn, user, keep = abi.fnptr(lib, "nqs_ansatz_log_amp"), a, None
    if w["marshall"]:
        keep = abi.MarshallWrapper(fn, a, w["Lx"], w["Ly"])
        fn, user = abi.fnptr(lib, "nqs_marshall_log_amp"), ctypes.addressof(keep)
    mv = {"flip": abi.MOVE_FLIP, "exchange": abi.MOVE_EXCHANGE,
          "learned": abi.MOVE_LEARNED}.get(move or args.move, w["move"])
    opt = abi.options_defaults(num_chains=w["chains"], move=mv, sz_zero_start=w["sz0"],
                               periodic=w["periodic"],
                               precision={"fp64": abi.FP64, "fp32": abi.FP32,
                                          "mixed": abi.MIXED}[args.precision],
                               gram={"direct": abi.GRAM_DIRECT, "factored": abi.GRAM_FACTORED,
                                     "auto": abi.GRAM_AUTO}[args.gram])
    if mv == abi.MOVE_LEARNED:
        # the learned proposal as the configuration states it: Adam at 1e-3,
        # retrained every 10 iterations on the last 16384 chain samples
        opt.ar_lr = LEARNED_LR
        opt.ar_train_every = 10
        opt.ar_train_window = 16384
        if getattr(args, "ar_local", None) is not None:
            opt.ar_local = args.ar_local
        if getattr(args, "ar_epochs", None) is not None:
            opt.ar_epochs = args.ar_epochs
        if getattr(args, "ar_batch", None) is not None:
            opt.ar_batch = args.ar_batch
    s = lib.nqs_hip_sampler_create_ex(N, ctypes.byref(cfg), ctypes.byref(opt), fn, user)
    if not s:
        raise RuntimeError("nqs_hip_sampler_create_ex failed")
    # which MinSR form the engine uses (the rule of NQS_HIP_GRAM_AUTO)
    B = w["chains"] * w["per_chain"]
    factored = args.gram in ("factored", "auto") and w["kind"] in (5, 6)
    gram_b = B * B * (N + 2 * w["M"]) if factored else B * B * P
    dual = args.gram == "auto" and B * P * P + P ** 3 / 3 < gram_b + B ** 3 / 3
    return dict(cfg=cfg, a=a, s=s, fn=fn, user=user, keep=keep, N=N, P=P, opt=opt,
                factored=factored and not dual, dual=dual, move=mv)


def step(lib, w, st, method=None, cfg=None):
    info = abi.SrInfo()
    fname = {"minsr": "nqs_sr_step_minsr_full", "cg": "nqs_sr_step_custom_full",
             "holo": "nqs_sr_step_holomorphic_full"}[method or w["method"]]
    rc = getattr(lib, fname)(ctypes.byref(cfg or st["cfg"]), w["Lx"], w["Ly"], st["a"], st["s"],
                             st["fn"], st["user"], None, None, ctypes.byref(info))
    if rc != 0:
        raise RuntimeError(f"{fname} returned {rc}")
    return info


def optimise(lib, w, st, budget):
    """`budget` iterations of the workload's energy optimiser from the
    current state (for C5 the reference's kagome recipe, holomorphic SR with
    lr 0.02 and shift 1e-3: the real-projected MinSR of the timed step cannot
    move the phases a non-stoquastic ground state needs).  Returns the
    trace's tail (last 50 iterations: mean, binned error) and its error
    against exact diagonalisation where that is known."""
    method = w.get("opt_method", w["method"])
    cfg = abi.NqsConfig.from_buffer_copy(st["cfg"])
    cfg.learning_rate = w.get("opt_lr", cfg.learning_rate)
    cfg.sr_diag_shift = w.get("opt_shift", cfg.sr_diag_shift)
    trace = []
    t0 = time.perf_counter()
    for i in range(budget):
        info = step(lib, w, st, method, cfg)
        trace.append(info.mean_energy)
        if (i + 1) % 50 == 0:
            log(f"[bench] optimise {i + 1}/{budget}: E = {info.mean_energy:.6f}")
    tail = np.array(trace[-50:])
    nbin = 5
    bins = tail[: len(tail) // nbin * nbin].reshape(nbin, -1).mean(axis=1)
    out = dict(optimizer={"holo": "holomorphic CG-SR", "minsr": "MinSR", "cg": "CG-SR"}[method],
               iterations=budget, lr=cfg.learning_rate, shift=cfg.sr_diag_shift,
               wall_s=time.perf_counter() - t0, tail_mean=float(tail.mean()),
               tail_err=float(bins.std(ddof=1) / np.sqrt(nbin)) if len(tail) >= nbin else None,
               trace_every_10=[float(x) for x in trace[::10]])
    if w["ed"] is not None:
        out["e_ed"] = w["ed"]
        out["delta_e_rel"] = (out["tail_mean"] - w["ed"]) / abs(w["ed"])
    if not np.all(np.isfinite(trace)):
        out["nonfinite_at"] = int(np.argmin(np.isfinite(trace)))
    return out


# kernel symbol (as rocprofv3 names it) behind each timing tag
SYMBOL = {"gemm_potrf": "nqs::k_gemm_big<1>", "gemm_chain": "nqs::k_gemm_lat",
          "multi_gemm_g": "nqs::(anonymous namespace)::k_mb_gemm_g",
          "multi_gemm_z": "nqs::(anonymous namespace)::k_mb_gemm_z",
          "multi_rows": "nqs::(anonymous namespace)::k_mb_rows",
          "multi_vec": "nqs::(anonymous namespace)::k_mb_finish + k_mb_update + k_mb_dir",
          "gemm": "nqs::k_gemm_big<0>",
          "potrf_panel": "nqs::k_panel2",
          "eloc": "nqs::k_eloc_terms<double, true>", "gen_O": "nqs::k_gen_O<double>",
          "colsum": "nqs::k_colsum_partial", "mh": "nqs::k_mh<double, true, 64, 4>",
          "trsv": "nqs::k_sbsv_fwd", "gram_x3": "nqs::k_gemm_x3"}


def traffic_bytes(workload, tag):
    """HBM bytes per dispatch of the kernel behind `tag`, from the committed
    rocprofv3 --pmc summary of this workload (counter values scaled by the
    copy-probe calibration), or None."""
    path = os.path.join(ROOT, "profiles", f"{workload}_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    d = json.load(open(path)).get("bytes_per_dispatch", {})
    v = d.get(SYMBOL.get(tag, ""))
    return None if v is None else v["fetch"] + v["write"]


TAGS = ["mh", "ar_mh", "forward", "eloc", "gen_O", "colsum", "gemm", "gemm_small", "gram_x3",
        "gemm_potrf", "gemm_chain", "potrf_panel",
        "trsv", "trsv_inv", "gram_f", "multi_prep", "multi_gemm_z", "multi_rows", "multi_gemm_g", "multi_vec"]


def potrf_widths():
    """Wide-panel width rules of the Cholesky, as the library reads them from
    NQS_HIP_POTRF_WIDTHS ("w1:r1,w2:r2,...,wk": w1 while more than r1 columns
    are left, ...)."""
    rules = []
    for part in os.environ.get("NQS_HIP_POTRF_WIDTHS", "512").split(","):
        if not part:
            continue
        w, _, above = part.partition(":")
        w = int(w)
        if w >= 128 and w % 128 == 0:
            rules.append((w, int(above) if above else 0))
    return rules or [(512, 0)]


def potrf_gemm_flops(B, part="rest"):
    """FLOP of the blocked Cholesky's GEMM updates of a B x B matrix (lower
    elements only): wide panels made of 128-column outer panels; inside a
    wide panel each outer panel's rank-128 update reaches the wide panel's
    later columns, then the next wide panel's columns and the rest of the
    trailing triangle take one update of the wide panel's rank.  part
    "rest": the rest-of-trailing-matrix updates (`gemm_potrf`); "chain": the
    updates inside the wide panels and of the next wide panel's columns
    (`gemm_chain`)."""
    rules = potrf_widths()

    def width(rem):
        for w, above in rules:
            if rem > above:
                return min(w, rem)
        return min(rules[-1][0], rem)

    def lower_cols(r0, w):   # elements of columns [r0, r0 + w), rows >= r0, lower part
        return w * (w + 1) // 2 + (B - r0 - w) * w

    def inside(k0, w):
        end = min(k0 + w, B)
        return sum(2 * 128 * lower_cols(j + 128, end - j - 128)
                   for j in range(k0, end, 128) if j + 128 < end)

    k0 = 0
    wd = width(B)
    chain, rest = inside(0, wd), 0
    while k0 + wd < B:
        r0 = k0 + wd
        w1 = width(B - r0)
        m = B - r0 - w1
        chain += 2 * wd * lower_cols(r0, w1) + inside(r0, w1)
        rest += 2 * wd * (m * (m + 1) // 2)
        k0, wd = r0, w1
    return chain if part == "chain" else rest


def kernel_work(tag, w, st, launches, iters):
    """Algorithmic work of one launch of `tag`: (bound, amount, unit).
    `launches` were counted over `iters` SR iterations."""
    B = w["chains"] * w["per_chain"]
    P, N, M = st["P"], st["N"], w["M"]
    per_iter = max(launches / max(iters, 1), 1.0)
    if tag == "gemm":
        kdim = 2 * M if w["kind"] == 6 else M
        if st["dual"]:
            # parameter-space MinSR: lower triangle of S = O_c^T O_c, P(P+1)/2
            # dot products of length B
            return "mfma", P * (P + 1) * B / per_iter, "FLOP"
        if st["factored"]:
            # factored Gram: lower triangle of U U^T, K = M (2M complex)
            return "mfma", B * (B + 1) * kdim / per_iter, "FLOP"
        # lower triangle of T = O_c O_c^T: B(B+1)/2 dot products of length P
        return "mfma", B * (B + 1) * P / per_iter, "FLOP"
    if tag == "gram_x3":
        # the mixed mode's Gram on the split-BF16 MFMA (three BF16 products
        # per algorithmic multiply-add): lower triangle of S = O_c^T O_c
        # (K = B) or of T = O_c O_c^T (K = P)
        if st["dual"]:
            return "mfma", P * (P + 1) * B / per_iter, "FLOP"
        return "mfma", B * (B + 1) * P / per_iter, "FLOP"
    # the Cholesky factors the P x P matrix in the parameter-space solve
    n_chol = P if st["dual"] else B
    if tag == "gemm_potrf":
        return "mfma", potrf_gemm_flops(n_chol, "rest") / per_iter, "FLOP"
    if tag == "gemm_chain":
        return "mfma", potrf_gemm_flops(n_chol, "chain") / per_iter, "FLOP"
    if tag == "potrf_panel":
        # per inner panel: factor + inverse of the 64x64 block, and the rows
        # below solved by a 64-wide product (and the preceding panel's
        # rank-64 update for every second panel); one launch factors the two
        # inner panels of a 128-column outer panel
        B = n_chol
        npan = (B + 63) // 64
        tot = sum(2 * 64 ** 3 / 3 + 2 * (B - 64 * (k + 1)) * 64 * 64 * (2 if k % 2 else 1)
                  for k in range(npan))
        return "mfma", tot / per_iter, "FLOP"
    if tag in ("multi_gemm_g", "multi_gemm_z"):
        # one problem: one QGT product per CG iteration, 2 (M+1)(N+1) B FLOP
        return "mfma", 2.0 * (M + 1) * (N + 1) * B, "FLOP"
    if tag == "multi_vec":
        return "hbm", 13 * (M + 1) * (N + 1) * 8, "B"
    if tag == "multi_rows":
        return "hbm", 3 * (M + 1) * B * 8, "B"
    if tag == "gen_O":
        return "hbm", B * P * 8 + B * N * 4 + B * M * 8 * (2 if w["kind"] == 6 else 1), "B"
    if tag == "colsum":
        return "hbm", B * P * 8, "B"
    return "hbm", B * N * 4, "B"


def gpu_run(args, w, move=None, budget=0):
    """One sampler leg: set up, thermalise, optimise `budget` iterations
    (untimed; the energy error is reported), then W warm-up and K timed SR
    iterations of the workload's step at that state."""
    lib = abi.load_library()
    st = setup(lib, w, args, move)
    mname = {abi.MOVE_FLIP: "flip", abi.MOVE_EXCHANGE: "exchange", abi.MOVE_LEARNED: "learned"}[st["move"]]
    log(f"[bench] workload {args.workload} ({mname} moves): N={st['N']} P={st['P']} samples/iter="
        f"{w['chains'] * w['per_chain']}")
    lib.nqs_sampler_thermalize(st["s"])
    opt = optimise(lib, w, st, budget) if budget > 0 else None
    if opt:
        log(f"[bench] optimised {budget} iterations ({opt['optimizer']}): tail E = "
            f"{opt['tail_mean']:.6f}" + (f", dE = {opt['delta_e_rel']:.3e}" if "delta_e_rel" in opt else ""))

    def profiled_step():
        lib.nqs_hip_kernel_timing(",".join(TAGS).encode())
        step(lib, w, st)
        out = {}
        for t in TAGS:
            c, ms = ctypes.c_longlong(0), ctypes.c_double(0)
            if lib.nqs_hip_kernel_stats(t.encode(), ctypes.byref(c), ctypes.byref(ms)) == 0 and c.value:
                out[t] = (c.value, ms.value)
        lib.nqs_hip_kernel_timing(None)
        return out

    # which kernel dominates the iteration: the split of a step after the
    # warm-up (first-touch costs of the work buffers are out of it), or of
    # the first step where later steps replay one graph (no per-kernel times)
    per = profiled_step()
    for i in range(args.warmup):
        step(lib, w, st)
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    per = profiled_step() or per
    dom = max(per, key=lambda t: per[t][1]) if per else "gemm"
    log("[bench] kernel split of one iteration (launches, ms): " +
        ", ".join(f"{t}={c}/{ms:.2f}" for t, (c, ms) in sorted(per.items(), key=lambda kv: -kv[1][1])))
    C, R = w["chains"], w["per_chain"]
    series = np.zeros((C, 0))
    buf = np.zeros(C * R)
    lib.nqs_hip_kernel_timing(dom.encode())
    energies, infos, stages = [], [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        info = step(lib, w, st)
        n = lib.nqs_hip_sampler_last_energies(st["s"], abi.ptr(buf), C * R)
        series = np.concatenate([series, buf[: C * n].reshape(C, n)], axis=1)
        energies.append(info.mean_energy)
        infos.append(info.as_dict())
        sm = np.zeros(7)
        if lib.nqs_hip_last_stage_ms(st["s"], abi.ptr(sm)) == 0:
            stages.append(sm.copy())
    elapsed = time.perf_counter() - t0
    c, ms = ctypes.c_longlong(0), ctypes.c_double(0)
    lib.nqs_hip_kernel_stats(dom.encode(), ctypes.byref(c), ctypes.byref(ms))
    lib.nqs_hip_kernel_timing(None)
    log(f"[bench] timed {args.steps} iterations in {elapsed:.3f} s")
    tau, tau_lo, tau_hi = tau_ci(series)
    n_total = series.shape[0] * series.shape[1]
    ess = n_total / (2.0 * tau)
    sample_s = sum(sm[0] for sm in stages) * 1e-3 if len(stages) == args.steps else None
    learned = None
    if st["move"] == abi.MOVE_LEARNED:
        acc, props, ntr, nll = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_int(), ctypes.c_double()
        lib.nqs_hip_sampler_learned_stats(st["s"], ctypes.byref(acc), ctypes.byref(props),
                                          ctypes.byref(ntr), ctypes.byref(nll))
        learned = dict(acceptance=acc.value, proposals=props.value, trainings=ntr.value,
                       last_nll=nll.value, adam_lr=LEARNED_LR)
    avg_ms = ms.value / max(c.value, 1)
    launches, iters = c.value, args.steps
    if not c.value:
        # launches replayed in a graph are not timed one by one: use the
        # launch-by-launch profiling iteration
        launches, ms_prof = per[dom]
        avg_ms, iters = ms_prof / launches, 1
    bound, work, unit = kernel_work(dom, w, st, launches, iters)
    if bound == "mfma":
        pk = PEAK_BF16 if dom == "gram_x3" else PEAK_FP64
        achieved, peak, runit = work / (avg_ms * 1e-3) / 1e12, pk / 1e12, "TFLOP/s"
    else:
        achieved, peak, runit = work / (avg_ms * 1e-3) / 1e9, PEAK_HBM / 1e9, "GB/s"
    res = dict(
        move=mname, ess_per_s=ess / sample_s if sample_s else ess / elapsed,
        ess_per_s_iteration=ess / elapsed,
        sample_ms=1e3 * sample_s / args.steps if sample_s else None, learned=learned,
        optimised=opt, ms_per_step=1e3 * elapsed / args.steps, tau=tau, tau_ci95=[tau_lo, tau_hi],
        samples_per_chain=int(series.shape[1]),
        energies=energies, acceptance=infos[-1]["acceptance_ratio"],
        stage_ms=np.mean(stages, axis=0).tolist() if stages else None,
        roofline=dict(kernel=dom, symbol=SYMBOL.get(dom), bound=bound, achieved=achieved, peak=peak,
                      unit=runit, frac=achieved / peak, traffic=traffic_bytes(args.workload, dom),
                      launches=c.value, avg_ms=avg_ms,
                      work_per_launch=work, work_unit=unit,
                      **({"mfma_products_per_flop": 3, "mfma_util": 3 * achieved / peak}
                         if dom == "gram_x3" else {})),
        split={t: dict(launches=cc, ms=mm) for t, (cc, mm) in per.items()}, N=st["N"], P=st["P"],
        dual=st["dual"], factored=st["factored"])
    lib.nqs_sampler_free(st["s"])
    lib.nqs_ansatz_free(st["a"])
    return res


MULTI_TAGS = ["mh", "forward", "eloc", "multi_prep", "multi_gemm_z", "multi_rows", "multi_gemm_g",
              "multi_vec"]


def multi_instances(w):
    rng = np.random.default_rng(w["seed"])
    Ns = rng.choice(np.array(w["sizes"]), w["n"])
    hs = rng.uniform(0.5, 1.5, w["n"])
    return [(int(N), float(h)) for N, h in zip(Ns, hs)]


def tfim_pbc_closed_form(N, h):
    k = (2 * np.arange(N) + 1) * np.pi / N
    return float(-np.sum(np.sqrt(1.0 + h * h - 2 * h * np.cos(k))))


def multi_gemm_flops(tag, inst, B, iters):
    """FLOPs of all launches of a batched GEMM tag over one step: the force
    product (two weight sets) and one product per CG iteration taken."""
    tot = 0.0
    for (N, _), it in zip(inst, iters):
        M1, N1 = 2 * N + 1, N + 1
        per = 2.0 * M1 * N1 * B
        tot += per * (2 + it) if tag == "multi_gemm_g" else per * it
    return tot


def multi_nsplit(N, B, target):
    """Split factor of the batched G product (the rule of the library's
    autotuned target: tiles x splits <= target, splits <= min(32, B/64))."""
    M1, N1 = 2 * N + 1, N + 1
    tiles = -(-M1 // 64) * -(-N1 // 64)
    max_split = 1
    while max_split * 2 <= 32 and max_split * 2 * 64 <= B:
        max_split *= 2
    ns = 1
    while ns * 2 <= max_split and tiles * ns * 2 <= target:
        ns *= 2
    return ns


def multi_bytes(tag, inst, B, target):
    """Algorithmic HBM bytes of one launch of a batched vector tag over all
    problems: multi_vec = finish (split partials, Om, p in; Ap out), update
    (x, p, r, Ap in; x, r out) and direction (r, p in; p out) of one CG
    iteration; multi_rows = [t, 1], Zt in, weighted [t, 1] out."""
    tot = 0
    for N, _ in inst:
        M1, N1 = 2 * N + 1, N + 1
        if tag == "multi_vec":
            tot += (multi_nsplit(N, B, target) + 12) * M1 * N1 * 8
        else:
            tot += 3 * M1 * B * 8
    return tot


def gpu_run_multi(args, w):
    lib = abi.load_library()
    inst = multi_instances(w)
    C, R = w["chains"], w["per_chain"]
    B = C * R
    log(f"[bench] workload {args.workload}: {len(inst)} instances, N = "
        f"{sorted(set(N for N, _ in inst))}, {B} samples each")
    probs, keep = [], []
    fn = abi.fnptr(lib, "nqs_ansatz_log_amp")
    for i, (N, h) in enumerate(inst):
        cfg = abi.config_defaults(ansatz=w["kind"], hamiltonian=w["ham"], transverse_field=h,
                                  rbm_hidden_units=2 * N, num_samples=B, num_thermalize=4 * N,
                                  num_decorrelate=N, rng_seed=1000 + i, learning_rate=w["lr"],
                                  sr_diag_shift=w["shift"], sr_cg_max_iters=100, sr_cg_tol=1e-6)
        a = lib.nqs_ansatz_create(ctypes.byref(cfg), N)
        P = lib.nqs_ansatz_num_params(a)
        par = np.ctypeslib.as_array(ctypes.cast(lib.nqs_ansatz_params_raw(a),
                                                ctypes.POINTER(ctypes.c_double)), (P,))
        par[:] = np.random.default_rng(1 + i).normal(0.0, 0.01, P)
        opt = abi.options_defaults(num_chains=C, periodic=1)
        s = lib.nqs_hip_sampler_create_ex(N, ctypes.byref(cfg), ctypes.byref(opt), fn, a)
        if not a or not s:
            raise RuntimeError("instance setup failed")
        probs.append(dict(cfg=cfg, a=a, s=s, N=N, h=h, P=P))
    n = len(probs)
    cfgs = (abi.NqsConfig * n)(*[p["cfg"] for p in probs])
    lx = (ctypes.c_int * n)(*[p["N"] for p in probs])
    ly = (ctypes.c_int * n)(*([1] * n))
    av = (ctypes.c_void_p * n)(*[p["a"] for p in probs])
    sv = (ctypes.c_void_p * n)(*[p["s"] for p in probs])
    m = lib.nqs_hip_multi_create(n, cfgs, lx, ly, av, sv)
    if not m:
        raise RuntimeError("nqs_hip_multi_create failed")
    for p in probs:
        lib.nqs_sampler_thermalize(p["s"])
    infos = (abi.SrInfo * n)()
    lib.nqs_hip_multi_step(m, infos)          # first step: autotune
    tg, us = (ctypes.c_int * 8)(), (ctypes.c_double * 8)()
    chosen = lib.nqs_hip_multi_tuning(m, tg, us, 8)
    ntune = sum(1 for i in range(8) if us[i] > 0)
    log("[bench] autotune (G tiles x splits target: CG product us; 0 = persistent launch): " +
        ", ".join(f"{tg[i]}: {us[i]:.1f}" for i in range(ntune)) + f" -> {chosen}")
    # one step launch by launch (CG as one launch per phase and iteration,
    # the timed configuration) for the per-kernel split, then one with the
    # CG in the persistent cooperative launch for comparison
    lib.nqs_hip_multi_use_graph(m, 0)
    lib.nqs_hip_multi_use_persistent(m, 0)
    lib.nqs_hip_kernel_timing(",".join(MULTI_TAGS).encode())
    lib.nqs_hip_multi_step(m, infos)
    per = {}
    for t in MULTI_TAGS:
        c, ms = ctypes.c_longlong(0), ctypes.c_double(0)
        if lib.nqs_hip_kernel_stats(t.encode(), ctypes.byref(c), ctypes.byref(ms)) == 0 and c.value:
            per[t] = (c.value, ms.value)
    lib.nqs_hip_kernel_timing(None)
    iters_seq = [infos[i].cg_iterations for i in range(n)]
    lib.nqs_hip_multi_use_persistent(m, 1)
    lib.nqs_hip_kernel_timing(",".join(MULTI_TAGS + ["multi_cg"]).encode())
    lib.nqs_hip_multi_step(m, infos)
    per_p = {}
    for t in MULTI_TAGS + ["multi_cg"]:
        c, ms = ctypes.c_longlong(0), ctypes.c_double(0)
        if lib.nqs_hip_kernel_stats(t.encode(), ctypes.byref(c), ctypes.byref(ms)) == 0 and c.value:
            per_p[t] = (c.value, ms.value)
    lib.nqs_hip_kernel_timing(None)
    lib.nqs_hip_multi_use_persistent(m, 0)
    lib.nqs_hip_multi_use_graph(m, 1)
    iters_prof = iters_seq
    iters_per = [infos[i].cg_iterations for i in range(n)]
    seq_iter_us = 1e3 * sum(per.get(t, (0, 0.0))[1] for t in
                            ("multi_gemm_z", "multi_rows", "multi_gemm_g", "multi_vec")) / max(iters_seq)
    cg_iter_us = 1e3 * per_p["multi_cg"][1] / max(iters_per) if "multi_cg" in per_p else None
    log(f"[bench] CG iteration of all problems: launch sequence {seq_iter_us:.1f} us (kernel time), "
        f"persistent launch {cg_iter_us if cg_iter_us is None else round(cg_iter_us, 1)} us")
    log("[bench] kernel split of one step (launches, ms): " +
        ", ".join(f"{t}={c}/{ms:.2f}" for t, (c, ms) in sorted(per.items(), key=lambda kv: -kv[1][1])))
    for i in range(args.warmup):
        lib.nqs_hip_multi_step(m, infos)
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    series = [np.zeros((C, 0)) for _ in range(n)]
    buf = np.zeros(B)
    t_total, t_sample = 0.0, 0.0
    smp = ctypes.c_double()
    for k in range(args.steps):
        t0 = time.perf_counter()
        lib.nqs_hip_multi_step(m, infos)
        t_total += time.perf_counter() - t0
        if lib.nqs_hip_multi_last_sample_ms(m, ctypes.byref(smp)) == 0:
            t_sample += smp.value * 1e-3
        for i in range(n):
            lib.nqs_hip_multi_last_energies(m, i, abi.ptr(buf), B)
            series[i] = np.concatenate([series[i], buf.reshape(R, C).T], axis=1)
    ms_step = 1e3 * t_total / args.steps
    log(f"[bench] timed {args.steps} steps: {ms_step:.3f} ms per step (all {n} instances)")
    taus = [pooled_tau(x) for x in series]
    ess_total = sum(x.size / (2.0 * t) for x, t in zip(series, taus))
    iters = [infos[i].cg_iterations for i in range(n)]
    energies = [infos[i].mean_energy for i in range(n)]
    conv = None
    conv_iters = args.budget or 0
    if conv_iters > 0:
        trace = np.zeros((n, conv_iters))
        for k in range(conv_iters):
            lib.nqs_hip_multi_step(m, infos)
            trace[:, k] = [infos[i].mean_energy for i in range(n)]
            if (k + 1) % 50 == 0:
                log(f"[bench] converge {k + 1}/{conv_iters}")
        tail = trace[:, -20:].mean(axis=1)
        e0 = np.array([tfim_pbc_closed_form(N, h) for N, h in inst])
        rel = (tail - e0) / np.abs(e0)
        conv = dict(iterations=conv_iters, delta_e_rel=[float(x) for x in rel],
                    delta_e_rel_max=float(rel.max()), delta_e_rel_median=float(np.median(rel)))
    dom = max(per, key=lambda t: per[t][1])
    c, tot_ms = per[dom]
    avg_ms = tot_ms / c
    if dom in ("multi_gemm_g", "multi_gemm_z"):
        work = multi_gemm_flops(dom, inst, B, iters_prof) / c
        achieved, peak, runit, bound = work / (avg_ms * 1e-3) / 1e12, PEAK_FP64 / 1e12, "TFLOP/s", "mfma"
    elif dom in ("multi_vec", "multi_rows"):
        work = multi_bytes(dom, inst, B, chosen)
        achieved, peak, runit, bound = work / (avg_ms * 1e-3) / 1e9, PEAK_HBM / 1e9, "GB/s", "hbm"
    else:
        # bytes the kernel must move: spins in, samples / energies out
        work = sum(B * N * 4 for N, _ in inst) / c
        achieved, peak, runit, bound = work / (avg_ms * 1e-3) / 1e9, PEAK_HBM / 1e9, "GB/s", "hbm"
    lib.nqs_hip_multi_free(m)
    for p in probs:
        lib.nqs_sampler_free(p["s"])
        lib.nqs_ansatz_free(p["a"])
    return dict(ess_per_s=ess_total / t_sample if t_sample > 0 else ess_total / t_total,
                ess_per_s_iteration=ess_total / t_total, sample_ms=1e3 * t_sample / args.steps,
                ms_per_step=ms_step, taus=taus,
                iters=iters, energies=energies, inst=inst, convergence=conv, chosen=chosen,
                tuning={int(tg[i]): float(us[i]) for i in range(ntune)},
                roofline=dict(kernel=dom, symbol=SYMBOL.get(dom), bound=bound, achieved=achieved,
                              peak=peak, unit=runit, frac=achieved / peak, traffic=None,
                              avg_ms=avg_ms),
                cg_iter_us=dict(launch_sequence=seq_iter_us, persistent_cooperative=cg_iter_us),
                split={t: dict(launches=cc, ms=mm) for t, (cc, mm) in per.items()},
                split_persistent={t: dict(launches=cc, ms=mm) for t, (cc, mm) in per_p.items()})


def cpu_baseline_multi(w, gpu):
    """Reference CPU path (1 core) on a bounded sample of each instance size:
    per-sample MH / E_loc / O cost on 8 samples per distinct N, and the CG
    MAC rate from one reference CG-SR step (N=32, Ns=256); the full step of
    all instances is extrapolated with the GPU run's CG iteration counts."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import load_oracle, load_reference
    lib, pre, kind = load_reference(), "nqs_", "reference"
    if lib is None:
        lib, pre, kind = load_oracle(), "orc_", "port"
    f = lambda nm: getattr(lib, pre + nm)
    fn = abi.fnptr(lib, pre + "ansatz_log_amp")
    B = w["chains"] * w["per_chain"]
    per_sample, per_sample_mh = {}, {}
    nb = 8
    for N in sorted(set(N for N, _ in gpu["inst"])):
        cfg = abi.config_defaults(ansatz=w["kind"], hamiltonian=0, rbm_hidden_units=2 * N,
                                  num_thermalize=0, num_decorrelate=N, rng_seed=7,
                                  rbm_init_scale=0.06)
        a = f("ansatz_create")(ctypes.byref(cfg), N)
        P = f("ansatz_num_params")(a)
        s = f("sampler_create")(N, ctypes.byref(cfg), fn, a)
        smp = np.zeros((nb, N), dtype=np.int32)
        t0 = time.perf_counter()
        f("sampler_batch")(s, nb, abi.ptr(smp))
        t_mh = (time.perf_counter() - t0) / nb
        e = np.zeros(nb)
        t0 = time.perf_counter()
        f("local_energy_batch")(ctypes.byref(cfg), N, 1, abi.ptr(smp), nb, fn, a, abi.ptr(e))
        # open chain: N - 1 bonds and N flips; periodic: N and N
        t_el = (time.perf_counter() - t0) / nb * (2 * N) / (2 * N - 1)
        g = np.zeros(P)
        t0 = time.perf_counter()
        for b in range(nb):
            f("ansatz_logpsi_gradient")(a, abi.ptr(np.ascontiguousarray(smp[b])), N, abi.ptr(g))
        t_o = (time.perf_counter() - t0) / nb
        per_sample[N] = t_mh + t_el + t_o
        per_sample_mh[N] = t_mh
        f("sampler_free")(s)
        f("ansatz_free")(a)
    # CG MAC rate: one reference CG-SR step, N = 32, Ns = 256
    N, ns = 32, 256
    cfg = abi.config_defaults(ansatz=w["kind"], hamiltonian=0, rbm_hidden_units=2 * N,
                              num_samples=ns, num_thermalize=0, num_decorrelate=N, rng_seed=9,
                              rbm_init_scale=0.06, sr_diag_shift=w["shift"], sr_cg_max_iters=100,
                              sr_cg_tol=1e-6)
    a = f("ansatz_create")(ctypes.byref(cfg), N)
    P = f("ansatz_num_params")(a)
    s = f("sampler_create")(N, ctypes.byref(cfg), fn, a)
    info = abi.SrInfo()
    t0 = time.perf_counter()
    f("sr_step")(ctypes.byref(cfg), N, 1, a, s, ctypes.byref(info))
    t_step = time.perf_counter() - t0
    macs = (info.cg_iterations * 2 + 2) * ns * P
    rate = macs / max(t_step - ns * per_sample[32] if 32 in per_sample else t_step * 0.5, 1e-6)
    f("sampler_free")(s)
    f("ansatz_free")(a)
    t_iter = 0.0
    for (N, _), it in zip(gpu["inst"], gpu["iters"]):
        Pn = 2 * N * N + 3 * N
        t_iter += B * per_sample[N] + (2 * it + 2) * B * Pn / rate
    ess_cpu = sum(B / (2.0 * t) for t in gpu["taus"]) / sum(B * per_sample_mh[N] for N, _ in gpu["inst"])
    ess_cpu_iter = sum(B / (2.0 * t) for t in gpu["taus"]) / t_iter
    sample = (f"{kind} CPU path, 1 core: {nb} samples per instance size of MH (N moves each), "
              f"E_loc (open twin scaled to the periodic bond count) and O rows, plus one "
              f"reference CG-SR step at N=32, Ns={ns} for the CG MAC rate "
              f"({rate / 1e9:.2f} GMAC/s); the step of all {len(gpu['inst'])} instances "
              f"({t_iter:.1f} s) is extrapolated with the GPU run's CG iteration counts and tau_int")
    return dict(value=ess_cpu, unit="effective samples/s", cores=1, kind=kind, sample=sample,
                sr_iter_s=t_iter, ess_per_s_iteration=ess_cpu_iter)


def cpu_baseline(w, gpu):
    """Reference CPU path on a bounded sample of the same workload, 1 core.

    Per-sample cost (MH moves, E_loc, O row) and the MinSR linear-algebra
    MAC rate are measured on the compiled reference (oracle/_ref) when it is
    present, else on the oracle port; the full-size iteration is
    extrapolated from exact operation counts."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import load_oracle, load_reference
    lib, pre, kind = load_reference(), "nqs_", "reference"
    if lib is None:
        lib, pre, kind = load_oracle(), "orc_", "port"
    f = lambda n: getattr(lib, pre + n)
    # the reference has open boundaries only: time the open twin and scale
    # E_loc by the ratio of off-diagonal bond counts
    N, Lx, Ly = gpu["N"], w["Lx"], w["Ly"]
    B_full = w["chains"] * w["per_chain"]
    cfg = abi.config_defaults(ansatz=w["kind"], hamiltonian=w["ham"], rbm_hidden_units=w["M"],
                              j2_coupling=0.5, num_thermalize=0, num_decorrelate=N, rng_seed=7,
                              rbm_init_scale=0.06)
    a = f("ansatz_create")(ctypes.byref(cfg), N)
    P = f("ansatz_num_params")(a)
    fn = abi.fnptr(lib, pre + "ansatz_log_amp")
    s = f("sampler_create")(N, ctypes.byref(cfg), fn, a)
    nb = 32
    smp = np.zeros((nb, N), dtype=np.int32)
    t0 = time.perf_counter()
    f("sampler_batch")(s, nb, abi.ptr(smp))
    t_mh = (time.perf_counter() - t0) / nb                       # N moves per sample
    if w["sz0"]:
        rng = np.random.default_rng(3)
        base = np.array([1] * (N // 2) + [-1] * (N - N // 2), dtype=np.int32)
        smp = np.ascontiguousarray(np.stack([rng.permutation(base) for _ in range(nb)]))
    e = np.zeros(nb)
    t0 = time.perf_counter()
    f("local_energy_batch")(ctypes.byref(cfg), Lx, Ly, abi.ptr(smp), nb, fn, a, abi.ptr(e))
    t_el_open = (time.perf_counter() - t0) / nb
    bonds_open = {0: (Lx - 1) * Ly + Lx * (Ly - 1) + N, 1: (Lx - 1) * Ly + Lx * (Ly - 1),
                  2: (Lx - 1) * Ly + Lx * (Ly - 1) + 2 * (Lx - 1) * (Ly - 1), 6: 6 * Lx * Ly}
    bonds_per = {0: 2 * N, 1: 2 * N, 2: 4 * N, 6: 6 * Lx * Ly}
    if w["Ly"] == 1:
        bonds_open[0], bonds_per[0] = (Lx - 1) + N, Lx + N
    t_el = t_el_open * bonds_per[w["ham"]] / bonds_open[w["ham"]]
    g = np.zeros(P)
    t0 = time.perf_counter()
    for b in range(nb):
        f("ansatz_logpsi_gradient")(a, abi.ptr(np.ascontiguousarray(smp[b])), N, abi.ptr(g))
    t_o = (time.perf_counter() - t0) / nb
    per_sample = t_mh + t_el + t_o
    # MinSR linear algebra rate: one full reference MinSR step at Ns = 768
    # (384 for the largest parameter counts)
    ns = 768 if P < 20000 else 384
    cfg2 = abi.config_defaults(ansatz=w["kind"], hamiltonian=w["ham"], rbm_hidden_units=w["M"],
                               j2_coupling=0.5, num_samples=ns, num_decorrelate=1, rng_seed=9,
                               rbm_init_scale=0.06, sr_diag_shift=1e-3)
    info = abi.SrInfo()
    f("sampler_thermalize")(s)
    t0 = time.perf_counter()
    f("sr_step_minsr_full")(ctypes.byref(cfg2), Lx, Ly, a, s, None, None, None, None, ctypes.byref(info))
    t_step = time.perf_counter() - t0
    per_sample_small = t_mh + t_el_open + t_o
    macs_small = ns * (ns + 1) / 2 * P + ns ** 3 / 6 + 2 * ns * P
    rate = macs_small / max(t_step - ns * per_sample_small, 1e-6)
    if w["method"] == "minsr":
        macs_full = B_full * (B_full + 1) / 2 * P + B_full ** 3 / 6 + 2 * B_full * P
    else:   # CG-SR: 100 iterations x 2 passes over O, plus means and force
        macs_full = 100 * 2 * B_full * P + 2 * B_full * P
    t_iter = B_full * per_sample + macs_full / rate
    f("sampler_free")(s)
    f("ansatz_free")(a)
    # the same definitions as the GPU line: effective samples per second of
    # the sampling phase (value) and of the whole iteration
    ess_cpu = (B_full / (2.0 * gpu["tau"])) / (B_full * t_mh)
    ess_cpu_iter = (B_full / (2.0 * gpu["tau"])) / t_iter
    sample = (f"{kind} CPU path, 1 core: {nb} samples of MH ({N} moves each), E_loc (open twin, "
              f"scaled to the periodic bond count) and O rows, plus one full MinSR step at "
              f"Ns={ns} for the linear-algebra MAC rate ({rate / 1e9:.2f} GMAC/s); the "
              f"{B_full}-sample iteration ({t_iter:.1f} s) is extrapolated from exact "
              f"operation counts with the GPU run's tau_int")
    return dict(value=ess_cpu, unit="effective samples/s", cores=1, kind=kind, sample=sample,
                sr_iter_s=t_iter, ess_per_s_iteration=ess_cpu_iter)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c5", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="mixed", choices=["fp64", "fp32", "mixed"])
    ap.add_argument("--gram", default="auto", choices=["direct", "factored", "auto"])
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the HBM copy probe on MIB mebibytes (counter calibration)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--c3", action="store_true",
                    help="C5: add the J1-J2 10x10 SR-iteration comparison with the reference CPU")
    ap.add_argument("--move", default=None, choices=["flip", "exchange", "learned"])
    ap.add_argument("--lr", type=float, default=None,
                    help="learning rate of the timed SR iterations (default: the workload's)")
    ap.add_argument("--budget", type=int, default=None, metavar="K",
                    help="optimisation iterations before timing (default: the workload's, C5 300)")
    ap.add_argument("--reference-records", default=None, metavar="DIR",
                    help="only write the reference benchmark's records (bench_nqs.c workloads, "
                         "bench_common.h schema) under DIR/nqs/")
    ap.add_argument("--ar-local", type=int, default=None, metavar="K",
                    help="learned proposal: exchange moves before each learned move")
    ap.add_argument("--ar-epochs", type=int, default=None, metavar="E",
                    help="learned proposal: passes over the window per retraining")
    ap.add_argument("--ar-batch", type=int, default=None, metavar="B",
                    help="learned proposal: training minibatch size (default 2048)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding of "
              f"this path is a design, not a build", file=sys.stderr)
        sys.exit(2)
    if args.reference_records:
        from spin_based_neural_network_amd.bench_records import write_records
        write_records(args.reference_records, log=log)
        return
    if args.copy_probe:
        lib = abi.load_library()
        nbytes = args.copy_probe << 20
        ms = lib.nqs_hip_copy_probe(nbytes, 10)
        print(json.dumps({"copy_probe_bytes": nbytes, "ms": ms, "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}))
        return
    w = WORKLOADS[args.workload]
    if w.get("multi"):
        return main_multi(args, w)
    budget = args.budget if args.budget is not None else w.get("budget", 0)
    if args.workload == "c5" and args.move is None:
        return main_c5(args, w, budget)
    gpu = gpu_run(args, w, None, budget)
    cpu = None
    if not args.no_cpu:
        log("[bench] timing the reference CPU path on a bounded sample")
        cpu = cpu_baseline(w, gpu)
    out = line(args, w, gpu, cpu)
    print(json.dumps(out))


def line(args, w, gpu, cpu):
    """The bench JSON line of one sampler leg."""
    out = {
        "metric": METRIC,
        "value": gpu["ess_per_s"],
        "unit": "effective samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gpu["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32+f64", "mixed": "f32+bf16+f16x3+f64"}[args.precision],
        "data": "synthetic (random-init parameters N(0,0.01), chains started at random Sz=0 states)",
        "config": {"workload": args.workload + ": " + w["desc"], "N": gpu["N"], "P": gpu["P"],
                   "samples_per_iter": w["chains"] * w["per_chain"], "chains": w["chains"],
                   "method": w["method"], "gram": args.gram, "precision": args.precision,
                   "move": gpu["move"]},
        "sr_iter_ms": gpu["ms_per_step"],
        "sample_ms": gpu["sample_ms"],
        "ess_per_s_iteration": gpu["ess_per_s_iteration"],
        "tau_int": gpu["tau"],
        "tau_int_ci95": gpu["tau_ci95"],
        "acceptance": gpu["acceptance"],
        "energy_last": gpu["energies"][-1],
        "stage_ms": dict(zip(["sample", "eloc", "O_reduce", "solve", "update"], gpu["stage_ms"][:5]))
        if gpu["stage_ms"] else None,
        "kernel_split_ms": {k: v["ms"] for k, v in gpu["split"].items()},
        "roofline": {k: gpu["roofline"][k] for k in ("bound", "achieved", "peak", "unit", "frac",
                                                     "traffic", "kernel", "symbol", "avg_ms",
                                                     "mfma_products_per_flop", "mfma_util")
                     if k in gpu["roofline"]},
        "cpu_baseline": cpu,
    }
    if gpu["learned"]:
        out["learned"] = gpu["learned"]
    if gpu["optimised"]:
        out["optimised"] = gpu["optimised"]
    if cpu:
        out["speedup_vs_cpu_iter"] = cpu["sr_iter_s"] / (gpu["ms_per_step"] * 1e-3)
    return out


def leg_summary(g):
    d = dict(ess_per_s=g["ess_per_s"], ess_per_s_iteration=g["ess_per_s_iteration"],
             sr_iter_ms=g["ms_per_step"], sample_ms=g["sample_ms"], tau_int=g["tau"],
             tau_int_ci95=g["tau_ci95"], samples_per_chain=g["samples_per_chain"],
             acceptance=g["acceptance"])
    if g["learned"]:
        d["learned"] = g["learned"]
    if g["optimised"]:
        o = g["optimised"]
        d["delta_e"] = {k: o.get(k) for k in ("optimizer", "iterations", "lr", "shift", "tail_mean",
                                               "tail_err", "e_ed", "delta_e_rel", "wall_s",
                                               "nonfinite_at")}
    return d


def main_c5(args, w, budget):
    """The metric's configuration: the same C5 system with plain exchange
    Metropolis and with the learned proposal.  Each leg optimises its own
    wavefunction for `budget` iterations (energy error vs exact
    diagonalisation reported), then times K MinSR iterations at that state;
    value = the learned leg's effective samples per second of the sampling
    phase, ms_per_step its SR-iteration time."""
    legs = {mv: gpu_run(args, w, mv, budget) for mv in ("exchange", "learned")}
    g = legs["learned"]
    cpu = None
    if not args.no_cpu:
        log("[bench] timing the reference CPU path on a bounded sample")
        cpu = cpu_baseline(w, g)
    out = line(args, w, g, cpu)
    out["data"] = (f"synthetic: random-init parameters N(0,0.01), chains from random Sz=0 states; "
                   f"each sampler leg first optimises its own wavefunction for {budget} iterations "
                   f"({w.get('opt_method', w['method'])} SR, untimed), then times {args.steps} "
                   f"MinSR iterations at that state")
    out["samplers"] = {mv: leg_summary(legs[mv]) for mv in legs}
    ex = legs["exchange"]
    out["learned_vs_exchange"] = dict(
        ess_per_s=g["ess_per_s"] / ex["ess_per_s"],
        ess_per_s_iteration=g["ess_per_s_iteration"] / ex["ess_per_s_iteration"],
        sr_iter_ms=g["ms_per_step"] / ex["ms_per_step"])
    if args.c3:
        out["c3_vs_reference_cpu"] = c3_comparison(args)
    print(json.dumps(out))


def c3_comparison(args):
    """J1-J2 10x10 (configuration 3): GPU SR-iteration time (plain exchange
    MH, MinSR on the 16384^2 Gram) against the reference CPU iteration on 1
    core, for the >= 50x target."""
    w3 = WORKLOADS["c3"]
    a3 = argparse.Namespace(**vars(args))
    a3.workload, a3.steps, a3.warmup, a3.move = "c3", 5, 2, None
    g3 = gpu_run(a3, w3, None, 0)
    out = dict(workload="c3: " + w3["desc"], sr_iter_ms=g3["ms_per_step"],
               roofline={k: g3["roofline"][k] for k in ("kernel", "achieved", "unit", "frac")})
    if not args.no_cpu:
        log("[bench] c3: timing the reference CPU path on a bounded sample")
        cpu = cpu_baseline(w3, g3)
        out.update(cpu_sr_iter_s=cpu["sr_iter_s"], cpu_kind=cpu["kind"], cpu_cores=cpu["cores"],
                   cpu_sample=cpu["sample"],
                   speedup_vs_cpu_iter=cpu["sr_iter_s"] / (g3["ms_per_step"] * 1e-3))
    return out


def main_multi(args, w):
    gpu = gpu_run_multi(args, w)
    cpu = None
    if not args.no_cpu:
        log("[bench] timing the reference CPU path on a bounded sample")
        cpu = cpu_baseline_multi(w, gpu)
    out = {
        "metric": METRIC,
        "value": gpu["ess_per_s"],
        "unit": "effective samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gpu["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (random-init parameters N(0,0.01), chains started at random states)",
        "config": {"workload": args.workload + ": " + w["desc"], "instances": len(gpu["inst"]),
                   "samples_per_instance": w["chains"] * w["per_chain"], "method": "cg",
                   "split_target": gpu["chosen"], "cg_launch": "launch sequence in the graph"},
        "sr_iter_ms": gpu["ms_per_step"],
        "sample_ms": gpu["sample_ms"],
        "ess_per_s_iteration": gpu["ess_per_s_iteration"],
        "cg_iterations": gpu["iters"],
        "autotune_us": gpu["tuning"],
        "kernel_split_ms": {k: v["ms"] for k, v in gpu["split"].items()},
        "kernel_split_persistent_ms": {k: v["ms"] for k, v in gpu["split_persistent"].items()},
        "cg_iter_us": gpu["cg_iter_us"],
        "roofline": gpu["roofline"],
        "cpu_baseline": cpu,
    }
    if gpu["convergence"]:
        out["convergence"] = gpu["convergence"]
    if cpu:
        out["speedup_vs_cpu_iter"] = cpu["sr_iter_s"] / (gpu["ms_per_step"] * 1e-3)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
