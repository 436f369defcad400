"""Benchmark records in the reference's schema, measured through this library.

The reference's NQS benchmark (benchmarks/bench_nqs.c) writes one JSON
record per measurement with the emitter of benchmarks/bench_common.h:42-87:

    {"suite": ..., "name": ..., "utc_epoch": <int>, "os": ..., "arch": ...,
     "hostname": ..., "metrics": {<key>: <number>, ...}}

under benchmarks/results/<suite>/<name>.json.  ``write_records(out_dir)``
runs the same workloads (same configurations, sizes and timed calls as
bench_nqs.c:17-148, main :186-291) through the C ABI of libnqs_hip.so and
writes records with the same suite, names and metric keys, so the
reference's trend tracking reads them unchanged.  The Kitaev-Heisenberg rows
(kh_*) are not written: that Hamiltonian is outside this library's device
path.  Numbers are printed to 9 significant digits as the C emitter's %.9g.
"""
import ctypes
import json
import os
import platform
import socket
import time

import numpy as np

from . import abi

TFIM, KAGOME = 0, 6
CRBM = 6


def _emit(out_dir, suite, name, metrics):
    """One record, in the field order of bench_common.h:56-69."""
    d = os.path.join(out_dir, suite)
    os.makedirs(d, exist_ok=True)
    u = platform.uname()
    lines = ["{", f'  "suite": "{suite}",', f'  "name": "{name}",',
             f'  "utc_epoch": {int(time.time())},', f'  "os": "{u.system}",',
             f'  "arch": "{u.machine}",', f'  "hostname": "{socket.gethostname()}",',
             '  "metrics": {']
    items = []
    for k, v in metrics:
        items.append(f'    "{k}": {v:d}' if isinstance(v, int) else f'    "{k}": {v:.9g}')
    lines.append(",\n".join(items))
    lines += ["  }", "}"]
    path = os.path.join(d, f"{name}.json")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    json.load(open(path))          # the record parses
    return path


def _kagome_cfg(samples, **kw):
    return abi.config_defaults(ansatz=CRBM, hamiltonian=KAGOME, kh_K=1.0, kh_J=1.0,
                               j_coupling=1.0, kagome_pbc=1, num_samples=samples,
                               num_thermalize=128, num_decorrelate=1, rbm_hidden_units=8, **kw)


def _sampler_rate(lib, cfg, N):
    a = lib.nqs_ansatz_create(ctypes.byref(cfg), N)
    s = lib.nqs_sampler_create(N, ctypes.byref(cfg), abi.fnptr(lib, "nqs_ansatz_log_amp"), a)
    if not a or not s:
        raise RuntimeError("ansatz / sampler creation failed")
    lib.nqs_sampler_thermalize(s)
    n = cfg.num_samples
    buf = np.zeros((n, N), dtype=np.int32)
    t0 = time.perf_counter()
    if lib.nqs_sampler_batch(s, n, abi.ptr(buf)) != 0:
        raise RuntimeError("nqs_sampler_batch failed")
    dt = time.perf_counter() - t0
    lib.nqs_sampler_free(s)
    lib.nqs_ansatz_free(a)
    return n / dt


def _sr_rate(lib, cfg, Lx, Ly, N, steps, holomorphic):
    a = lib.nqs_ansatz_create(ctypes.byref(cfg), N)
    fn = abi.fnptr(lib, "nqs_ansatz_log_amp")
    s = lib.nqs_sampler_create(N, ctypes.byref(cfg), fn, a)
    lib.nqs_sampler_thermalize(s)
    info = abi.SrInfo()
    t0 = time.perf_counter()
    for _ in range(steps):
        rc = (lib.nqs_sr_step_holomorphic(ctypes.byref(cfg), Lx, Ly, a, s, fn, a, ctypes.byref(info))
              if holomorphic else lib.nqs_sr_step(ctypes.byref(cfg), Lx, Ly, a, s, ctypes.byref(info)))
        if rc != 0:
            raise RuntimeError("SR step failed")
    dt = time.perf_counter() - t0
    lib.nqs_sampler_free(s)
    lib.nqs_ansatz_free(a)
    return steps / dt


def _local_energy_rate(lib, cfg, Lx, Ly, N):
    a = lib.nqs_ansatz_create(ctypes.byref(cfg), N)
    fn = abi.fnptr(lib, "nqs_ansatz_log_amp")
    s = lib.nqs_sampler_create(N, ctypes.byref(cfg), fn, a)
    lib.nqs_sampler_thermalize(s)
    n = cfg.num_samples
    batch = np.zeros((n, N), dtype=np.int32)
    lib.nqs_sampler_batch(s, n, abi.ptr(batch))
    re, im = np.zeros(n), np.zeros(n)
    t0 = time.perf_counter()
    lib.nqs_local_energy_batch_complex(ctypes.byref(cfg), Lx, Ly, abi.ptr(batch), n, fn, a,
                                       abi.ptr(re), abi.ptr(im))
    dt = time.perf_counter() - t0
    lib.nqs_sampler_free(s)
    lib.nqs_ansatz_free(a)
    return n / dt


def write_records(out_dir, lib=None, log=print):
    """Run the bench_nqs.c workloads and write their records; returns the
    paths written."""
    lib = lib or abi.load_library()
    paths = []
    for L, n in ((4, 4096), (6, 2048), (8, 1024)):          # bench_nqs.c:187-199
        cfg = abi.config_defaults(num_samples=n, num_thermalize=128, num_decorrelate=1)
        r = _sampler_rate(lib, cfg, L * L)
        paths.append(_emit(out_dir, "nqs", f"sampler_L{L}",
                           [("L", L), ("samples", n), ("samples_per_second", r)]))
        log(f"nqs sampler L={L}: {r:.1f} samples/sec")
    for L, n, steps in ((4, 256, 20), (6, 128, 10)):        # :201-216
        cfg = abi.config_defaults(hamiltonian=TFIM, transverse_field=1.0, num_samples=n,
                                  num_thermalize=128, num_decorrelate=1, learning_rate=1e-3,
                                  sr_diag_shift=1e-2, sr_cg_max_iters=20)
        r = _sr_rate(lib, cfg, L, L, L * L, steps, False)
        paths.append(_emit(out_dir, "nqs", f"sr_step_L{L}",
                           [("L", L), ("samples_per_step", n), ("steps_run", steps),
                            ("sr_steps_per_second", r)]))
        log(f"nqs SR step L={L}: {r:.2f} steps/sec")
    # kagome rows (:221-291); the kh_* rows need the Kitaev-Heisenberg kernel
    cfg = _kagome_cfg(2048)
    r = _local_energy_rate(lib, cfg, 2, 2, 12)
    paths.append(_emit(out_dir, "nqs", "kagome_local_energy_2x2",
                       [("Lx", 2), ("Ly", 2), ("num_samples", 2048), ("local_energy_per_second", r)]))
    log(f"nqs kagome_local_energy_2x2: {r:.1f} eval/sec")
    cfg = _kagome_cfg(2048)
    r = _sampler_rate(lib, cfg, 12)
    paths.append(_emit(out_dir, "nqs", "kagome_sampler_2x2",
                       [("Lx", 2), ("Ly", 2), ("samples", 2048), ("samples_per_second", r)]))
    log(f"nqs kagome_sampler_2x2: {r:.1f} samples/sec")
    cfg = _kagome_cfg(256, learning_rate=1e-3, sr_diag_shift=1e-2, sr_cg_max_iters=20,
                      sr_cg_tol=1e-7)
    r = _sr_rate(lib, cfg, 2, 2, 12, 10, True)
    paths.append(_emit(out_dir, "nqs", "kagome_sr_holomorphic_2x2",
                       [("Lx", 2), ("Ly", 2), ("samples_per_step", 256), ("steps_run", 10),
                        ("sr_steps_per_second", r)]))
    log(f"nqs kagome_sr_holomorphic_2x2: {r:.2f} steps/sec")
    return paths
