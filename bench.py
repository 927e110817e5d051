"""Throughput of the hot path on one MI355X: batched plain Langevin dynamics
(the headline), with C1, TEA-HI (C3), uneven-batch (C4) and MLP-HI (C5)
sections.

Headline workload (batched plain Langevin, FP32, no HI): 256 independent
trajectories of a 13-protofilament x 40-dimer microtubule (1040 beads each,
266,240 beads in total), per-trajectory HybridTaus streams, bottom ring fixed,
LJ pair list cutoff 8 nm rebuilt every 100 steps, synthetic lattice jittered
N(0, 0.05 nm).  One step = forces + Euler-Maruyama for every bead of the
batch, and the list rebuild every 100th step.  The warm-up is followed by as
many untimed steps as it takes to reach a rebuild step, so the timed window
of K steps starts with a rebuild and holds ceil(K / 100) of them.
value = monomer-steps/s (beads x steps / s) over that window.

Sections: "c1" the 312-bead single trajectory; "tea_hi" batched TEA
hydrodynamics (64 x 1040 beads) with its drift and covariance errors against
the FP64 Ermak-McCammon oracle; "c4" 48 uneven periodic trajectories with
free dimers, assembly, cell lists and TEA; "mlp_hi" device training of the
learned HI correction from the engine's own TEA runs and its deployment on
32 x 43,250-bead systems beside plain TEA.

Prints one JSON line.  Each section reports its dominant kernel's roofline
fraction (algorithmic bytes or flops per launch / mean launch time from HIP
events on the engine stream) and the CPU oracle timed on this host's cores on
a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

ALG_BYTES_PER_MONOMER_STEP = 112   # 6 x f32 state in + out (48 B) + two HybridTaus uint4 in + out (64 B)
HBM_PEAK_GBS = 8000.0
BF16_DENSE_PEAK_TFS = 2500.0


def build_workload(ntraj=256, n_dimers=40, seed=20240601):
    from mt_amd import lattice, params, topology
    p = params.derive(params.BASE)
    b = topology.replicate(lattice.microtubule(n_dimers), ntraj, 0.05, seed)
    b.coords = lattice.wrap_angles(b.coords)
    return p, b


def read_traffic(name):
    """HBM bytes per dispatch of the dominant kernel from the committed
    rocprofv3 --pmc summary: FETCH_SIZE and WRITE_SIZE (KiB per dispatch),
    each scaled by what the same counter reads for a plain 1 GiB copy kernel
    (the calibration factors recorded beside them)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        m = re.match(r"\s*(fetch_kib_per_dispatch|write_kib_per_dispatch|fetch_calibration|write_calibration)"
                     r"\s*[:=]\s*([0-9.eE+-]+)", line)
        if m:
            vals[m.group(1)] = float(m.group(2))
    if not {"fetch_kib_per_dispatch", "write_kib_per_dispatch"} <= vals.keys():
        return None
    return 1024.0 * (vals["fetch_kib_per_dispatch"] * vals.get("fetch_calibration", 1.0)
                     + vals["write_kib_per_dispatch"] * vals.get("write_calibration", 1.0))


TEA_KERNELS = {1: "k_tea_pairs_sym", 2: "k_tea_pairs_mfma"}   # tea_chunk -> pair kernel (else k_tea_pairs)


def tea_kernel(chunk):
    return TEA_KERNELS.get(chunk, "k_tea_pairs")


def read_tea_traffic(chunk):
    """HBM bytes per k_tea_pairs dispatch for j-chunks of `chunk` beads, from
    the committed rocprofv3 --pmc summary (FETCH_SIZE, WRITE_SIZE per chunk
    length, scaled by the copy-kernel calibration recorded there)."""
    path = os.path.join(ROOT, "profiles", "k_tea_pairs_c3.txt")
    if not os.path.exists(path):
        return None
    cal = {"fetch_calibration": 1.0, "write_calibration": 1.0}
    found = None
    for line in open(path):
        m = re.match(r"chunk (\d+): fetch_kib_per_dispatch = ([0-9.]+), write_kib_per_dispatch = ([0-9.]+)", line)
        if m and int(m.group(1)) == chunk:
            found = (float(m.group(2)), float(m.group(3)))
        m = re.match(r"(fetch_calibration|write_calibration) = ([0-9.]+)", line)
        if m:
            cal[m.group(1)] = float(m.group(2))
    if found is None:
        return None
    return 1024.0 * (found[0] * cal["fetch_calibration"] + found[1] * cal["write_calibration"])


def read_profile_keys(name, keys):
    """`key = value` lines of a committed profile summary (profiles/)."""
    path = os.path.join(ROOT, "profiles", name)
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        m = re.match(r"\s*([a-z_0-9]+)\s*=\s*([0-9.eE+-]+)\s*$", line)
        if m and m.group(1) in keys:
            out[m.group(1)] = float(m.group(2))
    return out


def cpu_baseline(ntraj_sample, steps_sample, n_dimers=40):
    """The FP32 CPU oracle (OpenMP over beads, all host cores) on a bounded
    sample: ntraj_sample trajectories of the same microtubule, steps 0 ..
    steps_sample - 1 (the rebuilds at multiples of 100 included)."""
    from oracle.oracle import Oracle, generate_seeds, max_threads
    p, b = build_workload(ntraj=ntraj_sample, n_dimers=n_dimers)
    o = Oracle(p, b, max_lj=64, precision="f32")
    seeds = generate_seeds(1234567, 2 * b.nb)
    t0 = time.perf_counter()
    rc = o.run(0, steps_sample, seeds=seeds)
    dt = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError(f"cpu oracle status {rc}")
    return {"value": b.nb * steps_sample / dt, "unit": "monomer-steps/s", "cores": max_threads(), "kind": "port",
            "seconds": dt,
            "sample": f"{ntraj_sample} x {b.nb // ntraj_sample}-bead trajectories x {steps_sample} steps (from "
                      f"step 0, rebuilds at multiples of 100), FP32 oracle, OpenMP"}


def plain_force_error(p, b, state, ntraj=4):
    """Forces of the engine's bead_force (the code every plain kernel runs)
    against the FP64 oracle at a state of the timed run, first ntraj
    trajectories: the largest |dF| over the largest |F| of the same kind
    (xyz, angles), and per bead |dF_i| / |F_i| (median, 99th percentile) over
    beads whose |F_i| is at least 1e-3 of the largest."""
    from mt_amd.engine import Engine, MODE_STREAMED
    from oracle.oracle import Oracle
    _, sub = build_workload(ntraj=ntraj)
    sub.coords = np.ascontiguousarray(state[:sub.nb])
    eng = Engine(p, sub.copy(), max_lj=64, mode=MODE_STREAMED)
    f = eng.forces().astype(np.float64)
    eng.close()
    orc = Oracle(p, sub.copy(), max_lj=64, precision="f64")
    orc.rebuild()
    fo = orc.forces()
    out = {"beads": int(sub.nb), "reference": "FP64 CPU oracle"}
    for name, cols in (("xyz", slice(0, 3)), ("angles", slice(3, 6))):
        d = np.abs(f[:, cols] - fo[:, cols])
        out[f"max_rel_{name}"] = float(d.max() / np.abs(fo[:, cols]).max())
        nrm = np.linalg.norm(fo[:, cols], axis=1)
        keep = nrm >= 1e-3 * nrm.max()
        per = np.linalg.norm(f[:, cols] - fo[:, cols], axis=1)[keep] / nrm[keep]
        out[f"per_bead_median_{name}"] = float(np.median(per))
        out[f"per_bead_p99_{name}"] = float(np.percentile(per, 99))
    return out


def c1_bench(steps, cpu=True):
    """C1: one 312-bead microtubule (13 PF x 12 dimers), the reference's
    CPU-runnable case, where launch latency rather than bandwidth bounds a
    step."""
    from mt_amd.engine import Engine, MODE_RESIDENT
    p, b = build_workload(ntraj=1, n_dimers=12, seed=1234567)
    eng = Engine(p, b, max_lj=48)
    tuned, report = eng.autotune(200)
    print("[bench] c1 autotune: " + "; ".join(report), file=sys.stderr, flush=True)
    eng.run(100)                     # warm-up, ends on a rebuild step
    t0 = time.perf_counter()
    eng.run(steps)
    elapsed = time.perf_counter() - t0
    resident = eng.mode == MODE_RESIDENT
    eng.close()
    out = {"value": b.nb * steps / elapsed, "unit": "monomer-steps/s", "steps": steps, "warmup": 100,
           "ms_per_step": elapsed * 1e3 / steps,
           "config": {"workload": "C1: 1 trajectory x 312-bead 13PFx12-dimer microtubule, LJ list 8 nm rebuilt "
                                  "every 100 steps, HybridTaus noise", "beads": b.nb,
                      "mode": "resident" if resident else "streamed", "tuning": tuned},
           "note": "one workgroup's work per step: latency-bound, no roofline applies"}
    if cpu:
        out["cpu_baseline"] = cpu_baseline(1, 200000, n_dimers=12)
        out["cpu_baseline"]["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def tea_error_vs_fp64():
    """Relative error of the GPU TEA drift against the FP64 exact (Cholesky)
    Ermak-McCammon drift on the 312-bead microtubule: one zero-noise step with
    dt = 1000 ps so float32 positions resolve the displacement."""
    from mt_amd import lattice, params, topology
    from mt_amd.engine import Engine, MODE_STREAMED
    from oracle import em_exact
    from oracle.oracle import Oracle
    p = params.derive(dict(params.TEA, tea_epsilon_freq=1, dt=1000.0))
    b = topology.replicate(lattice.microtubule(12), 1, 0.05, 1234567)
    orc = Oracle(p, b.copy())
    orc.rebuild()
    F = orc.forces()
    eng = Engine(p, b.copy(), mode=MODE_STREAMED)
    eng.set_noise(np.zeros((1, b.nb, 6), np.float32))
    eng.run(1)
    g = eng.get_state().astype(np.float64)
    eng.close()
    c0 = b.coords.astype(np.float64)
    mob = (b.fixed == 0) & (b.extra == 0)
    exact = em_exact.drift(c0[:, :3], F[:, :3], p.tea_a, p.dt, p.gammaR)
    d = g[:, :3] - c0[:, :3]
    return float(np.abs(d[mob] - exact[mob]).max() / np.abs(exact[mob]).max())


def tea_covariance_error(eng, b, t=0):
    """TEA's noise covariance against the exact one for trajectory t of a
    TEA engine, in FP64 from the engine's own last refresh (beta, ci):
    |B B^T - D_hat|_F / |D_hat|_F with D_hat = L L^T the normalised RPY
    tensor of the FP64 Cholesky Ermak-McCammon oracle, and the largest
    per-coordinate variance error |(B B^T)_kk - 1|."""
    from oracle import em_exact
    s0, s1 = int(b.off[t]), int(b.off[t + 1])
    pos = eng.get_state()[s0:s1, :3].astype(np.float64)
    _, beta = eng.tea_stats()
    ci = eng.tea_ci()[s0:s1]
    Dh = em_exact.rpy_hat(pos, eng.params.tea_a)
    B = em_exact.tea_b_from(Dh, beta[t], ci)
    return {"beads": s1 - s0, "frobenius_rel": em_exact.covariance_error(B, Dh),
            "variance_max_abs": float(np.abs((B * B).sum(1) - 1.0).max()), "beta": float(beta[t])}


def tea_bench(steps, warmup, cpu=True, chunk=0):
    """TEA-HI dynamics: 64 trajectories x 1040 beads, a = 2 nm, water
    viscosity (128 kcal ps / (mol nm^3)), dt = 1 ps, epsilon/beta every 100
    steps."""
    from mt_amd import lattice, params, topology
    from mt_amd.engine import Engine, MODE_STREAMED
    p = params.derive(params.TEA)
    b = topology.replicate(lattice.microtubule(40), 64, 0.05, 77)
    b.coords = lattice.wrap_angles(b.coords)
    eng = Engine(p, b, max_lj=48, mode=MODE_STREAMED)
    if chunk:
        # a forced pair kernel (profiling one of them), no autotune
        eng.set_tuning(tea_chunk=chunk)
        tuned, report = eng.tuning(), [f"tea_chunk forced to {chunk}"]
    else:
        tuned, report = eng.autotune(100)   # 100-step trials (~8 ms each): the shapes differ by ~10%
    print("[bench] tea autotune: " + "; ".join(report), file=sys.stderr, flush=True)
    eng.run(warmup)
    # timed region: graph-replayed step segments
    t0 = time.perf_counter()
    eng.run(steps)
    elapsed = time.perf_counter() - t0
    # kernel timing: the same steps launched one by one with HIP events around the pair kernel
    eng.reset_timing()
    eng.set_timing(True)
    eng.run(min(steps, 200))
    pair_ms, pair_n = eng.kernel_time("tea_pairs")
    eng.set_timing(False)
    cov_c3 = tea_covariance_error(eng, b)
    eng.close()
    sizes = b.sizes().astype(np.float64)
    flops = float(36.0 * np.sum(sizes * (sizes - 1)))      # 3B x 3B by 3B x 2 product per step
    avg = pair_ms / 1e3 / max(pair_n, 1)
    kname = tea_kernel(tuned["tea_chunk"])
    out = {"value": b.nb * steps / elapsed, "unit": "monomer-steps/s", "steps": steps, "warmup": warmup,
           "ms_per_step": elapsed * 1e3 / steps,
           "config": {"workload": "TEA-HI: 64 trajectories x 1040-bead microtubule, a 2 nm, 0.89 cP, dt 1 ps, "
                                  "epsilon/beta every 100 steps", "beads": b.nb, "tuning": tuned},
           "autotune": report,
           "roofline": {"bound": "mfma", "achieved": flops / avg / 1e12, "peak": BF16_DENSE_PEAK_TFS,
                        "unit": "TFLOP/s", "frac": flops / avg / 1e12 / BF16_DENSE_PEAK_TFS,
                        "traffic": read_tea_traffic(tuned["tea_chunk"]),
                        "kernel": kname, "avg_launch_ms": avg * 1e3, "alg_flops_per_launch": flops,
                        "note": "algorithmic work 36 B(B-1) flops per trajectory-step against the bf16 dense "
                                "MFMA peak; the kernel evaluates the RPY blocks and their products on the "
                                "vector ALUs in FP32 (see fp32_valu)"},
           "fp32_valu": {"achieved": flops / avg / 1e12, "peak": 157.3, "unit": "TFLOP/s",
                         "frac": flops / avg / 1e12 / 157.3,
                         "note": "the same flops against the FP32 vector peak (157.3 counts packed FMAs)"},
           "force_error_vs_fp64": tea_error_vs_fp64(),
           "covariance_error_vs_fp64_cholesky": {"c3_trajectory_0": cov_c3, "c1_312_beads": tea_c1_covariance()}}
    if cpu:
        from oracle.oracle import Oracle, generate_seeds, max_threads
        pc = params.derive(params.TEA)
        bc = topology.replicate(lattice.microtubule(40), 64, 0.05, 77)
        o = Oracle(pc, bc, max_lj=64, precision="f32")
        seeds = generate_seeds(1234567, 2 * bc.nb)
        ns = 100
        t0 = time.perf_counter()
        o.run(0, ns, seeds=seeds)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": bc.nb * ns / dt, "unit": "monomer-steps/s", "cores": max_threads(),
                               "kind": "port", "seconds": dt,
                               "sample": f"the same 64 x 1040-bead batch x {ns} TEA steps from step 0 (one "
                                         "epsilon/beta refresh, one list rebuild), FP32 oracle, all-pairs TEA, "
                                         "OpenMP"}
        out["cpu_baseline"]["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def tea_c1_covariance():
    """TEA covariance error on the 312-bead microtubule (C1 geometry) after
    one TEA step of the engine."""
    from mt_amd import lattice, params, topology
    from mt_amd.engine import Engine, MODE_STREAMED
    p = params.derive(params.TEA)
    b = topology.replicate(lattice.microtubule(12), 1, 0.05, 1234567)
    eng = Engine(p, b, max_lj=48, mode=MODE_STREAMED)
    eng.run(1)
    out = tea_covariance_error(eng, b)
    eng.close()
    return out


def c4_bench(steps, cpu=True):
    """C4: 48 uneven trajectories, a microtubule of 20-125 dimers per
    protofilament plus 2k-20k free dimers at 10-50 uM in a periodic box each,
    assembly (bond breaking and re-forming), cell-list rebuilds and TEA."""
    from mt_amd import params, workloads
    from mt_amd.engine import Engine, MODE_STREAMED
    p = params.derive(params.TEA)
    b = workloads.uneven(48, dimers=(20, 125), free=(2000, 20000), conc=(10, 50), seed=4242)
    eng = Engine(p, b, max_lj=48, mode=MODE_STREAMED)
    tuned, report = eng.autotune(3)
    print("[bench] c4 autotune: " + "; ".join(report), file=sys.stderr, flush=True)
    eng.run(100)            # warm-up, ends on an epsilon refresh and list rebuild step
    t0 = time.perf_counter()
    eng.run(steps)
    elapsed = time.perf_counter() - t0
    eng.reset_timing()
    eng.set_timing(True)
    eng.run(10)
    pair_ms, pair_n = eng.kernel_time("tea_pairs")
    eng.set_timing(False)
    eng.close()
    sizes = b.sizes().astype(np.float64)
    flops = float(36.0 * np.sum(sizes * (sizes - 1)))
    avg = pair_ms / 1e3 / max(pair_n, 1)
    out = {"value": b.nb * steps / elapsed, "unit": "monomer-steps/s", "steps": steps, "warmup": 100,
           "ms_per_step": elapsed * 1e3 / steps,
           "config": {"workload": "C4: 48 uneven periodic trajectories, 520-3250 microtubule beads + 4k-40k free "
                                  "beads each, assembly, cell lists, TEA a 2 nm, 0.89 cP, dt 1 ps",
                      "beads": int(b.nb), "largest": int(sizes.max()), "smallest": int(sizes.min()),
                      "tuning": tuned},
           "roofline": {"bound": "mfma", "achieved": flops / avg / 1e12, "peak": BF16_DENSE_PEAK_TFS,
                        "unit": "TFLOP/s", "frac": flops / avg / 1e12 / BF16_DENSE_PEAK_TFS, "traffic": None,
                        "kernel": tea_kernel(tuned["tea_chunk"]),
                        "avg_launch_ms": avg * 1e3, "alg_flops_per_launch": flops}}
    if cpu:
        from oracle.oracle import Oracle, generate_seeds, max_threads
        t = int(np.argsort(sizes)[len(sizes) // 2])          # the median trajectory
        s0, s1 = int(b.off[t]), int(b.off[t + 1])
        from mt_amd import topology
        one = b      # the engine copied the batch at creation; b still holds the initial state
        sub = topology.Batch(coords=one.coords[s0:s1].copy(), off=np.array([0, s1 - s0], np.int64),
                             harm_count=one.harm_count[s0:s1].copy(), harm=one.harm[s0:s1].copy(),
                             lon_count=one.lon_count[s0:s1].copy(), lon=one.lon[s0:s1].copy(),
                             lat_count=one.lat_count[s0:s1].copy(), lat=one.lat[s0:s1].copy(),
                             fixed=one.fixed[s0:s1].copy(), extra=one.extra[s0:s1].copy(),
                             mon_type=one.mon_type[s0:s1].copy(), gtp=one.gtp[s0:s1].copy(),
                             on_tub=one.on_tub[s0:s1].copy(), zs=one.zs[t:t + 1].copy(), box=one.box[t:t + 1].copy())
        o = Oracle(p, sub, max_lj=64, precision="f32")
        seeds = generate_seeds(1234567, 2 * sub.nb)
        ns = 2
        t0 = time.perf_counter()
        o.run(0, ns, seeds=seeds)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": sub.nb * ns / dt, "unit": "monomer-steps/s", "cores": max_threads(),
                               "kind": "port", "seconds": dt,
                               "sample": f"the median trajectory of the batch ({sub.nb} beads) x {ns} TEA steps from "
                                         "step 0 (one epsilon/beta refresh, one all-pairs list rebuild), FP32 "
                                         "oracle, OpenMP"}
        out["cpu_baseline"]["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


MLP_FLOPS_PER_BEAD = 2 * (96 * 128 + 2 * 128 * 128 + 128 * 9)   # 96-128-128-128-9 network, one bead
MLP_LOCAL_BYTES_PER_BEAD = 16 + 96 * 4   # own position in, features out


def mlp_bench(steps, warmup, iters):
    """MLP-HI: training data from the engine's own TEA runs, device training,
    then deployment on 32 trajectories of a 125-dimer microtubule (3250 beads)
    with 20,000 free dimers (40,000 beads) at 30 uM in a periodic box, beside
    plain TEA and free-draining (no HI) runs of the same length from the same
    seeds.  Reports throughput, the free-dimer centre-of-mass diffusion
    coefficient of each, the sampled error the judge saw, and the MFMA
    roofline of the fused inference + integrator kernel."""
    from mt_amd import params, workloads, mlp_hi
    from mt_amd.engine import Engine, HI_MLP, HI_DEFAULT, MODE_STREAMED
    pt = params.derive(dict(params.TEA))
    t0 = time.perf_counter()
    sets = [(workloads.plain(64, 40, seed=11), 5, 20),
            (workloads.uneven(8, dimers=(40, 60), free=(3000, 5000), conc=(20, 50), seed=3), 6, 20),
            (workloads.uneven(8, dimers=(100, 125), free=(15000, 20000), conc=(20, 40), seed=17), 6, 25)]
    Xs, Ys = [], []
    for bt, frames, every in sets:
        x, y, _ = mlp_hi.training_data(pt, bt, frames=frames, every=every)
        Xs.append(x)
        Ys.append(y)
    X, Y = np.concatenate(Xs), np.concatenate(Ys)
    data_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    w, losses = mlp_hi.train(pt, sets[1][0], X, Y, iters=iters, bsize=8192, lr=2e-3)
    train_s = time.perf_counter() - t0
    print(f"[bench] mlp training: {len(X)} samples, {iters} iterations, {train_s:.1f} s, "
          f"loss {losses[:5].mean():.4g} -> {losses[-20:].mean():.4g}", file=sys.stderr, flush=True)
    b = workloads.uneven(32, dimers=(125, 125), free=(20000, 20000), conc=(30, 30), seed=5)
    res = {}
    for mode in ("mlp", "none", "tea"):
        p = pt if mode != "none" else params.derive({**mlp_hi._cfg_of(pt), "tea_on": False})
        eng = Engine(p, b, max_lj=48, mode=MODE_STREAMED, rseed=1234567)
        tuned, report = eng.autotune(5)
        print(f"[bench] mlp section {mode} autotune: " + "; ".join(report), file=sys.stderr, flush=True)
        if mode == "mlp":
            eng.mlp_set_weights(w)
            eng.set_hi_mode(HI_MLP, check_every=100, tolerance=0.02)
        else:
            eng.set_hi_mode(HI_DEFAULT)
        eng.run(warmup)
        s0 = eng.get_state()
        ta = time.perf_counter()
        eng.run(steps)
        el = time.perf_counter() - ta
        s1 = eng.get_state()
        D, De = mlp_hi.com_diffusion(b, s0, s1, steps * pt.dt)
        r = {"value": b.nb * steps / el, "ms_per_step": el * 1e3 / steps, "D_nm2_per_ps": D, "D_stderr": De,
             "tuning": tuned}
        if mode == "mlp":
            err, fb, bad = eng.mlp_status()
            r.update(sampled_error_max=float(np.max(err)), fallback=bool(fb), bad_beads=int(bad))
            eng.reset_timing()
            eng.set_timing(True)
            eng.run(50)
            k_ms, k_n = eng.kernel_time("mlp_step")
            l_ms, l_n = eng.kernel_time("mlp_local")
            eng.set_timing(False)
            flops = float(MLP_FLOPS_PER_BEAD) * b.nb
            if k_n == 0 or l_n == 0:   # the judge handed the run back to TEA: no MLP-HI launches to time
                r["roofline"] = r["roofline_mfma"] = None
            else:
                avg = k_ms / 1e3 / k_n
                r["roofline_mfma"] = {"bound": "mfma", "achieved": flops / avg / 1e12, "peak": BF16_DENSE_PEAK_TFS,
                                      "unit": "TFLOP/s", "frac": flops / avg / 1e12 / BF16_DENSE_PEAK_TFS,
                                      "traffic": None, "kernel": "k_mlp_step", "avg_launch_ms": avg * 1e3,
                                      "alg_flops_per_launch": flops}
                lavg = l_ms / 1e3 / l_n
                lbytes = float(MLP_LOCAL_BYTES_PER_BEAD) * b.nb
                traffic = read_profile_keys("k_mlp_c5.txt", ("k_mlp_local_traffic_bytes",))
                r["roofline"] = {"bound": "hbm", "achieved": lbytes / lavg / 1e9, "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": lbytes / lavg / 1e9 / HBM_PEAK_GBS,
                                 "traffic": traffic.get("k_mlp_local_traffic_bytes"), "kernel": "k_mlp_local",
                                 "avg_launch_ms": lavg * 1e3, "alg_bytes_per_launch": lbytes,
                                 "note": "the dominant kernel of an MLP-HI step (32-nearest-neighbour features); "
                                         "bound by dependent gathers, not bandwidth: contract 400 B per bead "
                                         "(own position in, 32 x 3 float features out)"}
        eng.close()
        res[mode] = r
        print(f"[bench] mlp section {mode}: {r['value']:.4g} monomer-steps/s, D {D:.5g}", file=sys.stderr,
              flush=True)
    tea = res["tea"]
    diss = mlp_dissociation(w)
    return {"value": res["mlp"]["value"], "unit": "monomer-steps/s", "steps": steps, "warmup": warmup,
            "ms_per_step": res["mlp"]["ms_per_step"],
            "config": {"workload": "MLP-HI: 32 trajectories x (125-dimer microtubule, 3250 beads + 20,000 free "
                                   "dimers, 40,000 beads) at 30 uM, periodic box, a 2 nm, 0.89 cP, dt 1 ps; "
                                   "judge every 100 steps at 2%", "beads": b.nb},
            "speedup_vs_gpu_tea": res["mlp"]["value"] / tea["value"],
            "diffusion_error_vs_tea": res["mlp"]["D_nm2_per_ps"] / tea["D_nm2_per_ps"] - 1.0,
            "diffusion_error_no_hi_vs_tea": res["none"]["D_nm2_per_ps"] / tea["D_nm2_per_ps"] - 1.0,
            "runs": res,
            "dissociation": diss,
            "training": {"samples": int(len(X)), "iterations": iters, "batch": 8192, "data_s": data_s,
                         "train_s": train_s, "loss_start": float(losses[:5].mean()),
                         "loss_end": float(losses[-20:].mean())},
            "roofline": res["mlp"]["roofline"], "roofline_mfma": res["mlp"]["roofline_mfma"]}


def mlp_dissociation(w, ntraj=16, steps=5000):
    """Dissociation rate under MLP-HI against TEA and against no HI: 16 GDP
    microtubules of 312 beads with the weakened wells and temperature at which
    protofilaments peel within 1 us (D_lat 1.5, D_long 6.0, 600 K), the
    reference's viscosity and dt (200 ps), a 2 nm, 5000 steps from the same
    seeds.  The rate is the count of bonded-list entries (longitudinal and
    lateral) lost since step 0 per trajectory per us, from the engine's own
    pair rebuild (src/compute_cuda.cu:527-674; the count src/updater.cpp:154-227
    turns into mt_length).  The network is the one trained for C5: the HI
    tensor it predicts depends on the geometry only.  MLP-HI runs twice: as
    deployed (the judge samples the network's error every 100 steps and hands
    the run to TEA above 2%), and with the judge's tolerance at 100% so that
    the network drives the whole run; the largest sampled error is reported
    for both."""
    from mt_amd import lattice, params, topology, mlp_hi
    from mt_amd.engine import Engine, HI_MLP, HI_DEFAULT, MODE_STREAMED
    over = dict(D_lat=1.5, D_long=6.0, Temp=600.0)
    pt = params.derive(dict(params.BASE, tea_on=True, tea_a=2.0, tea_epsilon_freq=100, tea_capricious=True,
                            tea_epsmax=999.0, **over))
    pn = params.derive(dict(params.BASE, **over))
    b = topology.replicate(lattice.microtubule(12), ntraj, 0.05, 99)
    b.coords = lattice.wrap_angles(b.coords)
    b.gtp[:] = 0
    n = b.nb // ntraj
    out, lost = {}, {}
    for mode in ("tea", "mlp", "mlp_no_judge", "none"):
        eng = Engine(pn if mode == "none" else pt, b.copy(), max_lj=64, mode=MODE_STREAMED, rseed=4321)
        eng.rebuild()
        L0 = eng.lists()
        if mode.startswith("mlp"):
            eng.mlp_set_weights(w)
            eng.set_hi_mode(HI_MLP, check_every=100, tolerance=0.02 if mode == "mlp" else 1.0)
        elif mode == "tea":
            eng.set_hi_mode(HI_DEFAULT)
        eng.run(steps)
        eng.rebuild()
        L1 = eng.lists()
        k = ((L0["lat_count"] - L1["lat_count"]) + (L0["lon_count"] - L1["lon_count"])).reshape(ntraj, n).sum(1)
        t_us = steps * pt.dt * 1e-6
        lost[mode] = k.astype(np.float64) / t_us
        out[mode] = {"per_trajectory_per_us": float(lost[mode].mean()),
                     "stderr": float(lost[mode].std(ddof=1) / np.sqrt(ntraj))}
        if mode.startswith("mlp"):
            err, fb, bad = eng.mlp_status()
            out[mode].update(sampled_error_max=float(np.max(err)), fallback=bool(fb))
        eng.close()
    for mode in ("mlp", "mlp_no_judge", "none"):
        d = lost[mode] - lost["tea"]
        out[f"error_{mode}_vs_tea"] = float(d.mean() / lost["tea"].mean())
        out[f"error_{mode}_vs_tea_stderr"] = float(d.std(ddof=1) / np.sqrt(ntraj) / lost["tea"].mean())
    out["workload"] = (f"{ntraj} GDP microtubules x 312 beads, D_lat 1.5, D_long 6.0, 600 K, viscosity 2.85e4, "
                       f"dt 200 ps, a 2 nm, {steps} steps from the same seeds; bonded-list entries lost per "
                       "trajectory per us; MLP-HI as deployed (judge at 2%) and with the judge's tolerance at 100%")
    print(f"[bench] dissociation per trajectory per us: TEA {out['tea']['per_trajectory_per_us']:.1f}, "
          f"MLP-HI {out['mlp']['per_trajectory_per_us']:.1f} (network throughout "
          f"{out['mlp_no_judge']['per_trajectory_per_us']:.1f}), no HI {out['none']['per_trajectory_per_us']:.1f}",
          file=sys.stderr, flush=True)
    return out


def mlp_profile(steps):
    """The MLP-HI deployment kernels alone, for profiling: the C5 batch with
    zero weights (the network's work does not depend on their values) and
    the judge's tolerance raised so the run stays on MLP-HI; no training and
    no TEA leg.  Prints per-kernel HIP-event averages, not a bench line."""
    from mt_amd import params, workloads
    from mt_amd.engine import Engine, HI_MLP, MLP_NPARAM, MODE_STREAMED
    pt = params.derive(dict(params.TEA))
    b = workloads.uneven(32, dimers=(125, 125), free=(20000, 20000), conc=(30, 30), seed=5)
    eng = Engine(pt, b, max_lj=48, mode=MODE_STREAMED, rseed=1234567)
    eng.mlp_set_weights(np.zeros(MLP_NPARAM, np.float32))
    eng.set_hi_mode(HI_MLP, check_every=100, tolerance=1e9)
    eng.run(20)
    t0 = time.perf_counter()
    eng.run(steps)
    el = time.perf_counter() - t0
    eng.reset_timing()
    eng.set_timing(True)
    eng.run(min(steps, 50))
    k_ms, k_n = eng.kernel_time("mlp_step")
    l_ms, l_n = eng.kernel_time("mlp_local")
    eng.set_timing(False)
    eng.close()
    return {"workload": "MLP-HI deployment kernels, C5 batch (1,384,000 beads), zero weights, profiling only",
            "ms_per_step": el * 1e3 / steps, "steps": steps,
            "k_mlp_step_avg_ms": k_ms / max(k_n, 1), "k_mlp_local_avg_ms": l_ms / max(l_n, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--mode", choices=["auto", "streamed", "resident"], default="auto")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tea", action="store_true")
    ap.add_argument("--tea-steps", type=int, default=2000)
    ap.add_argument("--tea-chunk", type=int, default=0,
                    help="force the TEA pair kernel (64/128/256 one-sided, 1 symmetric, 2 MFMA) instead of autotuning")
    ap.add_argument("--no-mlp", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no mte_autotune)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain section (profiling the TEA or MLP section alone); the line then carries "
                         "no headline value")
    ap.add_argument("--mlp-steps", type=int, default=300)
    ap.add_argument("--mlp-profile", action="store_true",
                    help="run only the MLP-HI deployment kernels (zero weights, no training) for profiling")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--c1-steps", type=int, default=5000)
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--c4-steps", type=int, default=100)
    ap.add_argument("--mlp-iters", type=int, default=20000)
    ap.add_argument("--cpu-traj", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=2000)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 1 GiB device-to-device copy kernel (for scaling FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 4}))
        return
    if args.mlp_profile:
        print(json.dumps(mlp_profile(args.mlp_steps)))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(trajectory batches shard across GPUs without communication).", file=sys.stderr)
        sys.exit(2)

    from mt_amd.engine import Engine, MODE_AUTO, MODE_RESIDENT, MODE_STREAMED
    mode = {"auto": MODE_AUTO, "streamed": MODE_STREAMED, "resident": MODE_RESIDENT}[args.mode]
    if args.no_plain:
        out = {"metric": "monomer-steps/s/GPU (batched plain Langevin)", "value": None, "note": "plain section skipped"}
        if not args.no_c1:
            out["c1"] = c1_bench(args.c1_steps, cpu=not args.no_cpu)
        if not args.no_tea:
            out["tea_hi"] = tea_bench(args.tea_steps, 20, cpu=not args.no_cpu, chunk=args.tea_chunk)
        if not args.no_c4:
            out["c4"] = c4_bench(args.c4_steps, cpu=not args.no_cpu)
        if not args.no_mlp:
            out["mlp_hi"] = mlp_bench(args.mlp_steps, 20, args.mlp_iters)
        print(json.dumps(out))
        return
    p, b = build_workload()
    eng = Engine(p, b, max_lj=48, mode=mode)
    tuned = None
    if not args.no_tune:
        tuned, report = eng.autotune(200)
        print("[bench] autotune: " + "; ".join(report), file=sys.stderr, flush=True)
    freq = p.ljpairsupdatefreq
    align = (-args.warmup) % freq     # untimed steps after the warm-up up to the next rebuild step
    print(f"[bench] beads={b.nb} trajectories={b.ntraj} mode={eng.mode} warmup={args.warmup}+{align}",
          file=sys.stderr, flush=True)
    eng.run(args.warmup + align)
    assert eng.step % freq == 0
    eng.reset_timing()
    eng.set_timing(True)
    t0 = time.perf_counter()
    eng.run(args.steps)   # synchronous at return; starts with a list rebuild
    elapsed = time.perf_counter() - t0
    step_ms, step_n = eng.kernel_time("step")
    reb_ms, reb_n = eng.kernel_time("rebuild")
    eng.set_timing(False)
    st = eng.get_state()
    if not np.all(np.isfinite(st)):
        raise RuntimeError("non-finite state after the timed run")
    # steady state: 10 full rebuild intervals from the next rebuild step
    eng.run((-eng.step) % freq)
    t1 = time.perf_counter()
    eng.run(10 * freq)
    steady_s = time.perf_counter() - t1
    value = b.nb * args.steps / elapsed
    steps_per_launch = args.steps / max(step_n, 1)
    avg_launch_s = step_ms / 1e3 / max(step_n, 1)
    alg_bytes = ALG_BYTES_PER_MONOMER_STEP * b.nb * steps_per_launch
    achieved = alg_bytes / avg_launch_s / 1e9
    kname = "k_resident" if eng.mode == MODE_RESIDENT else "k_step_plain"
    traffic = read_traffic(f"traffic_{kname}_c2.txt")
    valu = read_profile_keys(f"{kname}_c2.txt", ("valu_per_bead_step", "valu_issue_frac", "valu_ms_per_dispatch",
                                                 "steps_per_dispatch"))
    out = {
        "metric": "monomer-steps/s/GPU (batched plain Langevin)",
        "value": value,
        "unit": "monomer-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched plain Langevin: 256 trajectories x 1040-bead 13PFx40-dimer microtubule, "
                               "LJ list 8 nm rebuilt every 100 steps, HybridTaus noise; the timed window starts "
                               "on a rebuild step",
                   "trajectories": b.ntraj, "beads_per_trajectory": int(b.sizes()[0]), "beads": b.nb,
                   "mode": "resident" if eng.mode == MODE_RESIDENT else "streamed", "tuning": tuned,
                   "alignment_steps": align},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kname, "launches": step_n, "avg_launch_ms": avg_launch_s * 1e3,
                     "alg_bytes_per_launch": alg_bytes,
                     "valu": dict(valu, source=f"profiles/{kname}_c2.txt",
                                  note="the kernel keeps the state in LDS and is bound by VALU issue; the HBM "
                                       "figure is the 112 B/monomer-step contract")},
        "rebuild": {"launches": reb_n, "avg_ms": reb_ms / max(reb_n, 1)},
        "steady_state": {"steps": 10 * freq, "rebuilds": 10, "value": b.nb * 10 * freq / steady_s,
                         "ms_per_step": steady_s * 1e3 / (10 * freq)},
        "force_error_vs_fp64": plain_force_error(p, b, st),
    }
    eng.close()
    if not args.no_c1:
        out["c1"] = c1_bench(args.c1_steps, cpu=not args.no_cpu)
    if not args.no_tea:
        out["tea_hi"] = tea_bench(args.tea_steps, 20, cpu=not args.no_cpu, chunk=args.tea_chunk)
    if not args.no_c4:
        out["c4"] = c4_bench(args.c4_steps, cpu=not args.no_cpu)
    if not args.no_mlp:
        out["mlp_hi"] = mlp_bench(args.mlp_steps, 20, args.mlp_iters)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_traj, args.cpu_steps)
        out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
