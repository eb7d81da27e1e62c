"""Benchmark: DF K+residual assembly elements/s on one MI355X.

Workload (one "step" = one full K + residual assembly into the reference CSR pattern):
unit cube of 24^3 p=4 Lagrange hexes (13,824 elements, 125 nodes, 375x375 local K,
912,673 nodes, 1,728,900,297 CSR non-zeros), 6x6x6 Gauss-Lobatto points, interior
nodes jittered +-0.03h (seed 0), Mooney-Rivlin (mu1=0.5, mu2=0.5, lamb=20), evaluated
at a twisted current configuration.  Inputs are resident in HBM before timing.

Printed: one JSON line with value = elements/s (whole job), the stiffness kernel's
roofline (FP64 MFMA, with its HBM view), and the CPU baseline (the C restatement of the
reference assembly timed on a bounded sample of the same mesh, over up to 16 host
threads).  The default
data path is the colour-ordered direct scatter (the stiffness kernel writes the CSR);
--no-direct times the element-block store + per-node gather instead.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
import argparse
import contextlib
import ctypes
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (matrix and vector)
HBM_PEAK_GBS = 8000.0


def stiffness_flops_per_element(npe, G):
    """FP64 flops of the block contraction for one element: every entry of the upper
    triangle of the symmetric 3npe x 3npe element matrix (diagonal included) is a
    3G-term dot product (2 flops per term).  The p=4 kernels execute this: the three
    off-diagonal npe x npe blocks in full, the three symmetric diagonal blocks as
    upper triangles of 16 x 16 tiles."""
    n = 3 * npe
    return n * (n + 1) // 2 * 2 * 3 * G


def sf_flops_per_element(p1, q):
    """Flops of the sum-factorised kernel for one element (nine blocks): per block the
    point tensor C (54 FMA per point), then per zeta slice the zeta contraction (9 x p1
    x q^2 outputs of q FMA), the eta contraction (4 types x p1^3 x q outputs, nine
    (m,n) terms of q FMA in all) and the beta contraction (p1 x p1^3 tasks of 4q
    products and p1 x 4q FMA)."""
    q3, n3 = q ** 3, p1 ** 3
    per_block = (q3 * 54 * 2 + p1 * (9 * p1 * q * q * q * 2) + p1 * (9 * p1 * p1 * p1 * q * q * 2)
                 + p1 * (p1 * n3 * (4 * q + p1 * 4 * q * 2)))
    return 9 * per_block


SF_KERNELS = {
    0: "stiffness_sf_mfma2_kernel<5,6> (sum-factorised: zeta on the FP64 VALU, eta and beta on FP64 MFMA, "
       "two types per reduction)",
    3: "stiffness_sf_kernel<5,6> (sum-factorised, all FP64 VALU; MI355X's FP64 vector peak equals its FP64 "
       "matrix peak)",
    5: "stiffness_sf_mfma_kernel<5,6> (sum-factorised: zeta on the FP64 VALU, eta and beta on FP64 MFMA, "
       "one type per reduction)",
}
SF_KERNELS[2] = SF_KERNELS[4] = SF_KERNELS[0]
DIRECT_KERNEL = ("stiffness_sf_mfma2_kernel<5,6,true> (sum-factorised: zeta on the FP64 VALU, eta and beta on "
                 "FP64 MFMA, two types per reduction; writes the CSR itself, one launch per element colour)")


def direct_scatter_bytes(nelem, npe, G, nnz, table_bytes):
    """Compulsory HBM bytes of the direct-scatter stiffness phase: the 54 C-tensor fields
    of every point read once, the scatter table read once, every CSR value written once,
    and each further contribution to an entry (9 nelem npe^2 - nnz of them) read back
    once before it is added."""
    return 8 * nelem * G * 54 + table_bytes + 8 * nnz + 8 * (9 * nelem * npe * npe - nnz)


def compulsory_bytes(nelem, npe, nnode, nnz):
    """Algorithmic HBM bytes of one K+T assembly (the per-unit figure of the path's
    measurement plan): every CSR value written once (8 nnz), T written once (8 x 3
    nnode), X and x read once (2 x 8 x 3 nnode) and the int32 connectivity read once
    (4 nelem npe)."""
    return 8 * nnz + 8 * 3 * nnode + 16 * 3 * nnode + 4 * nelem * npe


PMC_SUMMARY = "assembly_hex_p4_24cube_pmc.md"
PMC_SOURCE = ("profiles/" + PMC_SUMMARY + ": rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, of the committed bench.py "
              "on this workload (copy-calibrated, KiB per dispatch summed over one assembly's launches)")


def copy_calibration(torch, mib):
    """One plain device-to-device copy of mib MiB: the known-size dispatch that scales the
    PMC byte counters of a profiled run."""
    a = torch.ones(mib << 17, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print("copy calibration: {} bytes read, {} bytes written".format(a.numel() * 8, a.numel() * 8),
          file=sys.stderr, flush=True)


def device_copy_gbs(torch, nbytes, reps=10):
    """Read+write rate of a plain device-to-device copy of nbytes (GB/s, bytes read plus
    bytes written), timed with HIP events."""
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(2):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2 * nbytes / (ms / 1e3) / 1e9


@contextlib.contextmanager
def no_gc():
    """A timed region without Python's cyclic collector: a full collection with torch
    loaded takes tens of milliseconds and would stall the launches of short loops."""
    gc.collect()
    gc.disable()
    try:
        yield
    finally:
        gc.enable()


def gather_bytes(nelem, npe, nnode, nnz, interior_nodes=0):
    """Compulsory HBM bytes of the gather: element blocks and tractions read once, CSR
    values and T written once -- less, when the stiffness kernel writes the rows of
    element-interior nodes itself, those nodes' 9 npe block entries read and CSR entries
    written and their 3 traction entries read and T entries written."""
    total = 8 * (nelem * 9 * npe * npe + nelem * 3 * npe + nnz + 3 * nnode)
    return total - 8 * interior_nodes * 2 * (9 * npe + 3)


def dense_b_flops_per_element(npe, G):
    """Work of the reference's dense-B algorithm per element (dgemm B.(H.B^T) per point
    plus the geometric, kinematic, residual and material terms)."""
    n, V = 3 * npe, 6
    return G * (2 * V * n * n + 2 * V * V * n + 2 * n * n + 6 * npe * npe + 108 * npe + 13 * n + 300)


def measured_traffic(summary, phase):
    """HBM bytes per launch of one phase's kernels: FETCH_SIZE + WRITE_SIZE from the
    committed rocprofv3 PMC summary (copy-calibrated), or None if there is none."""
    import re
    try:
        with open(os.path.join(ROOT, "profiles", summary)) as f:
            m = re.search(phase + r" traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read())
    except OSError:
        return None
    return float(m.group(1)) if m else None


def cpu_baseline(mesh, fs, mat, x, nsample, max_threads=16, one_core_sample=192):
    """Time oracle/liboracle_df.so (C restatement of _GlobalAssemblyDF_) on bounded
    samples of the mesh: on one core (the reference kernel is serial) over the first
    ``one_core_sample`` elements, and on every core of this process's affinity set (at
    most ``max_threads``: the CPU share of one GPU on an 8-GPU MI355X node is 16 cores,
    whatever the machine reports) over the first ``nsample`` elements split into equal chunks, each thread
    assembling its chunk into the chunk's own CSR pattern (the library call releases the
    GIL and keeps no global state) -- the reference's element-partitioned pool model.
    value = elements / wall time of the parallel section."""
    one = _cpu_run(mesh, fs, mat, x, one_core_sample, 1)
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    threads = max(1, min(max_threads, ncpu))
    allc = _cpu_run(mesh, fs, mat, x, max(nsample, 16 * threads), threads)
    if one is None or allc is None:
        return None
    allc["affinity_cores"] = ncpu
    allc["one_core"] = {"value": one["value"], "cores": 1, "sample": one["sample"]}
    allc["speedup_all_cores_vs_one"] = round(allc["value"] / one["value"], 2)
    return allc


def _cpu_run(mesh, fs, mat, x, nsample, threads):
    import concurrent.futures as cf
    from oracle import df_numpy as O
    so = os.path.join(ROOT, "oracle", "liboracle_df.so")
    if not os.path.exists(so):
        return None
    lib = ctypes.CDLL(so)
    lib.oracle_df_assemble.restype = ctypes.c_int
    threads = max(1, min(threads, nsample))
    chunk = nsample // threads
    prm = np.ascontiguousarray(mat.device_parameters())
    Jm, w = np.ascontiguousarray(fs.Jm), np.ascontiguousarray(fs.AllGauss.ravel())
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    tasks = []
    for i in range(threads):
        el = np.asarray(mesh.elements, dtype=np.int64)[i * chunk:(i + 1) * chunk]
        nodes, inv = np.unique(el, return_inverse=True)
        sub = np.ascontiguousarray(inv.reshape(el.shape).astype(np.int64))
        indptr, indices = O.sparsity_pattern(sub, nodes.size)
        tasks.append((sub, np.ascontiguousarray(np.asarray(mesh.points)[nodes]),
                      np.ascontiguousarray(np.asarray(x)[nodes]), indptr, indices, np.zeros(indices.size),
                      np.zeros(3 * nodes.size)))

    def run(t):
        sub, pts, xs, indptr, indices, V, T = t
        t0 = time.perf_counter()
        rc = lib.oracle_df_assemble(P(pts), P(sub), P(xs), P(Jm), P(w), ctypes.c_int64(w.size),
                                    ctypes.c_int64(sub.shape[1]), ctypes.c_int(mat.material_id), P(prm),
                                    ctypes.c_int(1), P(indptr), P(indices), P(V), P(T), ctypes.c_int64(0),
                                    ctypes.c_int64(sub.shape[0]))
        return rc, time.perf_counter() - t0

    with cf.ThreadPoolExecutor(threads) as ex:
        t0 = time.perf_counter()
        res = list(ex.map(run, tasks))
        dt = time.perf_counter() - t0
    if any(rc != 0 for rc, _ in res):
        return None
    per_core = chunk / np.mean([d for _, d in res])
    return {"value": threads * chunk / dt, "unit": "elements/s", "cores": threads, "kind": "port",
            "per_core": round(float(per_core), 2),
            "sample": "first {} elements of the same p=4 hex mesh in {} chunk(s) of {} on {} host thread(s), each "
                      "chunk K+T into its own CSR pattern, {:.1f} s wall, {:.1f} s of CPU; C restatement of "
                      "_GlobalAssemblyDF_ (dense B(HB^T) per point), gcc -O3 -march=native".format(
                          threads * chunk, threads, chunk, threads, dt, sum(d for _, d in res))}


def bench_batch(args):
    """256 seeded meshes (tet p1-4, hex p1-3, cube/cylinder, 200-20000 elements, random
    Neo-Hookean) assembled as one captured HIP graph; per-mesh check against assembling
    each mesh alone."""
    import torch
    from florence_amd import _lib, autotune
    from florence_amd.assembly import DeviceMeshData, assemble_device
    from florence_amd.batch import BatchedAssembly, random_batch, _np_class

    t0 = time.perf_counter()
    batch = random_batch(args.nmesh, seed=args.seed)
    t_mesh = time.perf_counter() - t0
    print("generated {} meshes, {} elements in {:.1f}s".format(len(batch), sum(m.nelem for m, _, _ in batch), t_mesh),
          file=sys.stderr, flush=True)
    ba = BatchedAssembly(batch, precision=args.precision)
    disp = lambda X: 0.02 * np.sin(np.pi * X[:, [1, 2, 0]]) + 0.01 * X
    ba.set_state(disp)
    print("{} groups, nnz {}".format(len(ba.groups), ba.nnz), file=sys.stderr, flush=True)
    cache = autotune.tune_batch(ba, cache=autotune.load_cache())
    autotune.save_cache(cache)
    print("tuned: {}".format({k: v["variant"] for k, v in cache.items()}), file=sys.stderr, flush=True)
    ba.assemble()
    torch.cuda.synchronize()
    # per-group phase times of one eager pass (library HIP events)
    _lib.profile_enable(True)
    ba.assemble()
    torch.cuda.synchronize()
    calls = _lib.profile_read()
    _lib.profile_enable(False)
    group_phases = {}
    for g, c in zip(ba.groups, [c for c in calls if c[0] == 1]):
        k = "{}-p{}".format(*g["key"])
        prev = group_phases.get(k, [0.0] * 4)
        group_phases[k] = [round(a + b, 3) for a, b in zip(prev, c[1])]
    print("group phases (qp, stiffness, traction, gather) ms: {}".format(group_phases), file=sys.stderr, flush=True)
    ba.capture()
    for _ in range(args.warmup):
        ba.replay()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with no_gc():
        ev0.record()
        for _ in range(args.steps):
            ba.replay()
        ev1.record()
        torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    print("batched graph: {:.3f} ms per batch".format(ms), file=sys.stderr, flush=True)
    # eager per-mesh reference: each mesh alone through the ordinary path, same variant
    t_check = time.perf_counter()
    worst, nbit, eager_ms = 0.0, 0, 0.0
    variants = {g["key"]: (g["variant"], g.get("sum_factorised", True), g.get("direct", True)) for g in ba.groups}
    for i, (m, key, mat) in enumerate(batch):
        g = next(gg for gg in ba.groups if i in gg["mesh_ids"])
        dm = DeviceMeshData(m, g["fs"])
        _lib.load().fl_set_tuning(_lib.FL_TUNE_STIFF_VARIANT_BASE + _np_class(dm.npe), variants[key][0] + 1)
        _lib.load().fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 2 if variants[key][1] else 1)
        X = np.asarray(m.points)
        x = X + disp(X)     # the same expression as set_state, so the same coordinates bitwise
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        assemble_device(dm, mat, x, precision=args.precision, direct=variants[key][2])
        e0.record()
        V, T, _ = assemble_device(dm, mat, x, precision=args.precision, direct=variants[key][2])
        e1.record()
        torch.cuda.synchronize()
        eager_ms += e0.elapsed_time(e1)
        ip, ix, data, Tb = ba.mesh_result(i)
        assert np.array_equal(ip, dm.indptr.cpu().numpy()) and np.array_equal(ix, dm.indices.cpu().numpy())
        Vh, Th = V.cpu().numpy(), T.cpu().numpy()
        err = max(np.abs(data - Vh).max() / np.abs(Vh).max(), np.abs(Tb - Th).max() / max(np.abs(Th).max(), 1e-300))
        worst = max(worst, err)
        nbit += int(np.array_equal(data, Vh) and np.array_equal(Tb, Th))
        del dm, V, T
        if i % 32 == 31:
            print("checked {} meshes, worst rel err {:.3e}".format(i + 1, worst), file=sys.stderr, flush=True)
    t_check = time.perf_counter() - t_check
    out = {"metric": "batched K+residual assembly elements/s ({} meshes, one graph launch)".format(len(batch)),
           "value": round(ba.nelem / (ms / 1e3), 1), "unit": "elements/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
           "config": {"workload": "{} seeded cube/cylinder meshes, tet p1-4 + hex p1-3, log-uniform 200-20000 "
                                  "elements, Neo-Hookean mu~U[0.5,2], lamb~U[5,50]".format(len(batch)),
                      "nelem": ba.nelem, "groups": len(ba.groups), "seed": args.seed},
           "per_mesh_eager_ms_sum": round(eager_ms, 3),
           "speedup_vs_per_mesh_calls": round(eager_ms / ms, 2),
           "tuned_variants": {"{}-p{}".format(*g["key"]): (("sum-factorised, direct scatter" if g.get("direct")
                                                             and g["dm"].scatter is not None else "sum-factorised")
                                                            if g.get("sum_factorised") and
                                                            g["dm"].tensor_basis is not None
                                                            else "dense v{}".format(g["variant"]))
                              for g in ba.groups},
           "group_phase_ms": group_phases,
           "tuning": cache,
           "per_mesh_check": {"max_rel_err": worst, "bitwise_equal_meshes": nbit, "meshes": len(batch),
                              "seconds": round(t_check, 1)},
           "status": ba.statuses()}
    print(json.dumps(out), flush=True)


def bench_learned(args):
    """Learned warm start: record trajectories of the twist/stretch family with the
    device Newton, train the MLP on the device, then solve held-out members with each
    start mode and tangent precision and report wall time to FP64 residual 1e-8."""
    import torch
    from florence_amd.learned import TwistStretchFamily, LockstepNewton, WarmStartMLP, build_dataset, train

    log = lambda s: print(s, file=sys.stderr, flush=True)
    fam = TwistStretchFamily(n=args.learn_n, nincr=args.nincr)
    t0 = time.perf_counter()
    if args.weights:
        net = WarmStartMLP.load(args.weights)
        gen_info = {"weights": os.path.basename(args.weights)}
    else:
        runs, iters = [], []
        for c in range(0, args.ntraj, args.traj_batch):
            seeds = range(c, min(c + args.traj_batch, args.ntraj))
            ln = LockstepNewton(fam, seeds)
            r = ln.solve("extrap", record=True)
            if r["failed"]:
                raise RuntimeError("trajectory batch {} failed".format(c))
            runs.append((ln.X.clone(), ln.free.clone(), r["trajectory"]))
            iters.append(sum(r["lockstep_iterations"]))
            log("trajectories {}..{}: {:.1f}s, {} lockstep Newton iterations".format(
                seeds[0], seeds[-1], r["wall_s"], iters[-1]))
            del ln
        t_gen = time.perf_counter() - t0
        net = WarmStartMLP(seed=args.seed)
        net.base_order = args.base_order
        F, T = build_dataset(net, fam, runs)
        del runs
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        hist = train(net, F, T, epochs=args.epochs, batch=8192, lr=2e-3, seed=args.seed, log=log)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t1
        gen_info = {"trajectories": args.ntraj, "increments": args.nincr, "rows": int(F.shape[0]),
                    "generate_s": round(t_gen, 1), "train_s": round(t_train, 1), "epochs": args.epochs,
                    "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1]}
        del F, T
        if args.save_weights:
            net.save(args.save_weights, **{k: v for k, v in gen_info.items() if not isinstance(v, str)})
    held = range(10 ** 6, 10 ** 6 + args.heldout)
    # (tangent precision, start, linear-solve forcing, CSR storage); the first is the
    # plain FP64 path
    modes = [("fp64", "previous", "fixed", "fp64"), ("fp64", "extrap", "fixed", "fp64"),
             ("fp64", "poly", "fixed", "fp64"), ("fp64", "learned", "fixed", "fp64"),
             ("fp64", "extrap", "adaptive", "fp64"), ("fp64", "poly", "adaptive", "fp64"),
             ("fp64", "learned", "adaptive", "fp64"), ("fp32", "poly", "adaptive", "fp64"),
             ("fp32", "learned", "adaptive", "fp64"), ("fp32", "poly", "adaptive", "fp32"),
             ("fp32", "learned", "adaptive", "fp32"), ("bf16", "learned", "adaptive", "fp32")]
    results, ref = {}, None
    for prec, start, forcing, csr in modes:
        ln = LockstepNewton(fam, held, precision=prec, forcing=forcing, csr=csr)
        ln.solve(start, predictor=net)           # warm-up (first-touch, kernel attributes)
        with no_gc():
            r = ln.solve(start, predictor=net, record=True)
        final = r["trajectory"][-1]
        if ref is None:
            ref = final
        key = "{}+{}+{}".format(prec, start, forcing) + ("+csr32" if csr == "fp32" else "")
        results[key] = {"wall_s": round(r["wall_s"], 3), "linear_s": round(r["linear_s"], 3),
                        "assembly_s": round(r["assembly_s"], 3), "residual_s": round(r["residual_s"], 3),
                        "predict_s": round(r["predict_s"], 4),
                        "lockstep_iterations": int(sum(r["lockstep_iterations"])),
                        "member_iterations_mean": float(r["member_iterations"].sum(0).mean()),
                        "pcg_iterations": int(sum(r["pcg_iterations"])),
                        "start_residual_mean": float(np.mean(r["start_residual"])) if r["start_residual"] else None,
                        "learned_chosen": int(sum(r["chose_learned"])), "failed": r["failed"],
                        "lockstep_per_increment": r["lockstep_iterations"],
                        "worst_residual_history_k10": [float("{:.3e}".format(v)) for v in r["residual_history"][9]]
                        if len(r["residual_history"]) > 9 else None,
                        "max_rel_diff_vs_fp64_previous": float(np.abs(final - ref).max() / np.abs(ref).max())}
        log("{}: {}".format(key, results[key]))
        del ln
    plain = results["fp64+previous+fixed"]["wall_s"]
    best_key = min((k for k in results if "learned" in k and "bf16" not in k and "fp32" in k),
                   key=lambda k: results[k]["wall_s"])
    out = {"metric": "Newton wall time to FP64 residual 1e-8 (held-out twist/stretch members, lockstep)",
           "value": results[best_key]["wall_s"], "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "higher_is_better": False, "vs_baseline": None, "dtype": best_key.split("+")[0], "data": "synthetic",
           "modes_key": "tangent precision + start + linear-solve forcing (+csr32: CSR values stored FP32)",
           "config": {"workload": "hex p=4 unit cubes {0}^3 (jitter 0.03h), MR(0.5,0.5,20), 6^3 GLL, x=0 clamped, "
                                  "x=1 twist U[0,60] deg + stretch U[0.8,1.5], {1} increments; {2} held-out members "
                                  "solved in lockstep".format(args.learn_n, args.nincr, args.heldout)},
           "best_mode": best_key, "speedup_vs_plain_fp64": round(plain / results[best_key]["wall_s"], 3),
           "modes": results, "training": gen_info}
    print(json.dumps(out), flush=True)


WEIGHTS = os.path.join(ROOT, "florence_amd", "data", "warmstart_twist_hex_p4.npz")


def newton_comparison(args, log):
    """The Newton half of the metric, on a workload that fits the default run: held-out
    members of the learned family (hex p=4 unit cubes of 2^3 cells, jitter 0.03h,
    MR(0.5,0.5,20), 6^3 GLL points, x=0 clamped, x=1 twist U[0,60] deg + stretch
    U[0.8,1.5], 30 increments) solved in lockstep to |R_free,i| <= 1e-8 R0_i per member.
    Plain FP64 path: FP64 tangent and CSR, the reference's predictor start.  Fast path:
    FP32 tangent, FP32-stored CSR, the committed learned warm start; beside it the same
    fast path started from the polynomial extrapolation (no network: per member the
    linear, quadratic or cubic one, whichever has the smallest FP64 residual).  All arms use the
    SAME linear solver and the same tolerance policy, for every policy listed."""
    import torch
    from florence_amd.learned import LockstepNewton, TwistStretchFamily, WarmStartMLP
    fam = TwistStretchFamily(n=2, nincr=30)
    net = WarmStartMLP.load(WEIGHTS)
    seeds = range(10 ** 6, 10 ** 6 + args.newton_members)
    arms = (("plain_fp64", "fp64", "previous", "fp64"), ("mixed_learned", "fp32", "learned", "fp32"),
            ("mixed_polynomial", "fp32", "poly", "fp32"))
    out = {"workload": "{} held-out members (seeds 1e6..) of the learned family: unit cubes 2^3 hex p=4, jitter "
                       "0.03h, MR(0.5,0.5,20), 6^3 GLL, x=0 clamped, x=1 twist U[0,60] deg + stretch U[0.8,1.5], "
                       "30 increments, lockstep, |R_free,i| <= 1e-8 R0_i".format(args.newton_members),
           "plain_fp64": "FP64 tangent, FP64 CSR, reference predictor start",
           "mixed_learned": "FP32 tangent, FP32-stored CSR, learned warm start (FP64 residual arbitrates)",
           "mixed_polynomial": "FP32 tangent, FP32-stored CSR, polynomial extrapolation of the last converged "
                               "states (order 1-3 per member, FP64 residual picks)",
           "policies": {}}
    for linear, forcing in (("mg", "fixed"), ("mg", "adaptive"), ("bj", "adaptive")):
        pol = {}
        for name, prec, start, csr in arms:
            ln = LockstepNewton(fam, seeds, precision=prec, forcing=forcing, csr=csr, linear=linear)
            ln.solve(start, predictor=net)           # warm-up (first-touch, kernel attributes)
            with no_gc():
                r = ln.solve(start, predictor=net, record=True)
            pol[name] = {"wall_s": round(r["wall_s"], 3), "linear_s": round(r["linear_s"], 3),
                         "assembly_s": round(r["assembly_s"], 3), "newton_iterations_per_increment":
                         r["lockstep_iterations"], "newton_iterations": int(sum(r["lockstep_iterations"])),
                         "linear_iterations": int(sum(r["pcg_iterations"])), "failed": r["failed"],
                         "mean_start_residual": float(np.mean(r["start_residual"])) if r["start_residual"] else None,
                         "final": r["trajectory"][-1]}
            del ln
            torch.cuda.empty_cache()
        a, b, c = pol["plain_fp64"], pol["mixed_learned"], pol["mixed_polynomial"]
        fa, fb, fc = a.pop("final"), b.pop("final"), c.pop("final")
        key = "{}+{}".format("multigrid" if linear == "mg" else "block_jacobi", forcing)
        pol["speedup"] = round(a["wall_s"] / b["wall_s"], 3)
        pol["speedup_polynomial"] = round(a["wall_s"] / c["wall_s"], 3)
        pol["max_rel_diff_displacement"] = max(float(np.abs(fb - fa).max() / np.abs(fa).max()),
                                               float(np.abs(fc - fa).max() / np.abs(fa).max()))
        out["policies"][key] = pol
        log("newton {}: plain {:.3f}s ({} its), mixed+learned {:.3f}s ({} its), {:.2f}x; mixed+polynomial "
            "{:.3f}s ({} its), {:.2f}x".format(key, a["wall_s"], a["newton_iterations"], b["wall_s"],
                                               b["newton_iterations"], pol["speedup"], c["wall_s"],
                                               c["newton_iterations"], pol["speedup_polynomial"]))
    return out


NEWTON_MODES = {"fixed": ("fp64", "previous", "fixed", "fp64"), "adaptive": ("fp64", "previous", "adaptive", "fp64"),
                "fast": ("fp32", "poly", "adaptive", "fp32")}


def twisted_cube_newton(n, linear, lin_rtol, log, which=("fixed", "adaptive", "fast")):
    """The twisted-cube solve on one mesh: hex p=4 unit cube n^3 (jitter 0.03h, seed 0),
    MR(0.5, 0.5, 20), 6^3 Gauss-Lobatto points, x=0 clamped, x=1 twisted 30 degrees, 10
    increments, each to |R_free| <= 1e-8 R0 -- three modes of the lockstep driver:
    the plain FP64 path (FP64 tangent and CSR, the reference's predictor start) with
    fixed linear tolerance (lin_rtol) and with the adaptive one, and the fast path (FP32
    tangent, FP32-stored CSR, the polynomial extrapolation of the last converged states
    as start -- per member the order, 1 to 3, with the smallest FP64 residual -- adaptive
    linear tolerance).  The
    multigrid's coarse levels are per-mesh set-up, like the sparsity pattern: built
    before the timed solve, their time is mg_setup_s."""
    import torch
    from florence_amd.learned import LockstepNewton, TwistStretchFamily
    fam = TwistStretchFamily(n=n, nincr=10, theta_range=(30.0, 30.0), stretch_range=(1.0, 1.0))
    modes, ref = {}, None
    for prec, start, forcing, csr in [NEWTON_MODES[k] for k in which]:
        ln = LockstepNewton(fam, [0], precision=prec, forcing=forcing, csr=csr, lin_rtol=lin_rtol, linear=linear)
        log("lockstep {} {} {} csr {}: {} dofs, nnz {}".format(prec, start, forcing, csr, 3 * ln.dm.nnode, ln.dm.nnz))
        with no_gc():
            r = ln.solve(start, record=True, log=log)
        final = r["trajectory"][-1]
        ref = final if ref is None else ref
        key = "{}+{}+{}".format(prec, start, forcing) + ("+csr32" if csr == "fp32" else "")
        modes[key] = {"wall_s": round(r["wall_s"], 3), "linear_s": round(r["linear_s"], 3),
                      "assembly_s": round(r["assembly_s"], 3), "residual_s": round(r["residual_s"], 3),
                      "mg_setup_s": round(ln.setup_s, 3),
                      "newton_iterations": r["lockstep_iterations"], "pcg_iterations_total":
                      int(sum(r["pcg_iterations"])), "linear_iterations": list(r["pcg_iterations"]),
                      "failed": r["failed"],
                      "max_rel_diff_vs_plain": float(np.abs(final - ref).max() / np.abs(ref).max())}
        log("{}: wall {:.3f}s, Newton {}, linear iterations {}".format(key, modes[key]["wall_s"],
                                                                      modes[key]["newton_iterations"],
                                                                      modes[key]["pcg_iterations_total"]))
        del ln
        torch.cuda.empty_cache()
    return modes


def bench_newton(args):
    """The twisted-cube solve: hex p=4 unit cube n^3 (jitter 0.03h, seed 0), MR(0.5, 0.5,
    20), 6^3 Gauss-Lobatto points, x=0 clamped, x=1 twisted 30 degrees, 10 increments.

    * For n <= 8 (``--femsolver``): FEMSolver with the reference's Newton semantics
      (tolerance 1e-8 on the re-based relative residual), FP64 and the mixed-precision
      tangent.
    * Always: the lockstep driver on this one mesh with a fixed criterion
      |R_free| <= 1e-8 R0 -- the plain FP64 path (the reference's predictor start,
      linear solves to 1e-10) against the fast path (FP32 tangent, FP32-stored CSR,
      extrapolated start, adaptive linear tolerance)."""
    import torch

    log = lambda m: print(m, file=sys.stderr, flush=True)
    out_modes = {}
    if args.femsolver:
        from florence_amd.boundary_condition import BoundaryCondition
        from florence_amd.solver import FEMSolver, LinearSolver
        from florence_amd.workloads import clamp_and_twist, cube_hex_p4
        mesh, fs, mat = cube_hex_p4(args.n)

        class Form(object):
            ndim = nvar = 3
            fields = "mechanics"
            function_spaces = (fs, None, None)

        ref = None
        for prec in ("fp64", "fp32"):
            bc = BoundaryCondition()
            bc.SetDirichletCriteria(clamp_and_twist, mesh, 30.0)
            if args.linear == "mg":
                from florence_amd.multigrid import MultigridSolver
                lin = MultigridSolver(rtol=args.lin_rtol, maxit=5000)
            else:
                lin = LinearSolver(rtol=args.lin_rtol, maxit=500000, check_every=25)
            solver = FEMSolver(optimise=True, number_of_load_increments=10, newton_raphson_tolerance=1e-8,
                               linear_solver=lin, tangent_precision=prec,
                               csr_precision="fp32" if prec != "fp64" else "fp64", report_log_level=1)
            with no_gc():
                res = solver.Solve(formulation=Form(), mesh=mesh, material=mat, boundary_condition=bc)
            final = res.TotalDisp[:, :, -1].cpu().numpy()
            ref = final if ref is None else ref
            key = "femsolver+" + prec
            out_modes[key] = {"wall_s": round(res.wall_time, 3), "assembly_s": round(res.assembly_time, 3),
                              "newton_iterations": list(res.newton_iterations),
                              "linear_iterations": list(res.cg_iterations),
                              "cg_iterations_total": int(sum(res.cg_iterations)),
                              "failed": bool(res.newton_raphson_failed_to_converge),
                              "max_rel_diff_vs_fp64": float(np.abs(final - ref).max() / np.abs(ref).max())}
            log("{}: {}".format(key, out_modes[key]))
            del solver, res
            torch.cuda.empty_cache()
    if args.copy_mib:
        copy_calibration(torch, args.copy_mib)
    modes = tuple(args.modes.split(","))
    out_modes.update(twisted_cube_newton(args.n, args.linear, args.lin_rtol, log, modes))
    plain = out_modes.get("fp64+previous+adaptive", {}).get("wall_s")
    fast = out_modes.get("fp32+poly+adaptive+csr32", {}).get("wall_s")
    out = {"metric": "Newton wall time to FP64 residual 1e-8 (twisted hex p=4 cube)", "value": fast, "unit": "s",
           "n_gpus": 1, "steps": 1, "warmup": 0, "higher_is_better": False, "vs_baseline": None, "dtype": "fp32",
           "data": "synthetic", "speedup_vs_plain_fp64": round(plain / fast, 3) if plain and fast else None,
           "config": {"workload": "unit cube {0}^3 hex p=4 (jitter 0.03h), 6^3 GLL, MR(0.5,0.5,20), x=0 clamped, "
                                  "x=1 twisted 30 deg, 10 increments, {1}".format(
                                      args.n, "p-multigrid flexible CG" if args.linear == "mg" else
                                      "block-Jacobi PCG"),
                      "lin_rtol": args.lin_rtol, "linear": args.linear},
           "modes": out_modes}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=24, help="cells per cube edge")
    ap.add_argument("--cpu-sample", type=int, default=1024,
                    help="elements in the CPU baseline sample, split over up to 16 host threads (0 = skip)")
    ap.add_argument("--stiffness-minw", type=int, default=0, help="stiffness kernel variant (waves/SIMD), 0 = default")
    ap.add_argument("--stiffness-variant", type=int, default=-1, help="p=4 stiffness tile variant (-1 = default)")
    ap.add_argument("--dense", action="store_true", help="dense MFMA stiffness even where sum factorisation applies")
    ap.add_argument("--gather-form", type=int, default=0, choices=[0, 1, 2, 3],
                    help="gather form bits: 1 one element at a time, 2 no single-element kernel")
    ap.add_argument("--no-direct", action="store_true",
                    help="store element blocks and gather them per node instead of the colour-ordered direct "
                         "scatter from the stiffness kernel")
    ap.add_argument("--sf-form", type=int, default=0, choices=[0, 3, 5],
                    help="sum-factorised stiffness form: 3 all-VALU, 5 MFMA one type per reduction (0: the "
                         "library default, MFMA two types per reduction)")
    ap.add_argument("--workload", default="assembly", choices=["assembly", "batch", "learned", "newton"],
                    help="assembly: the headline K+T metric; batch: the batched-meshes form; "
                         "learned: warm-start training and the Newton time comparison")
    ap.add_argument("--learn-n", type=int, default=2, help="learned: cells per cube edge of the family")
    ap.add_argument("--nincr", type=int, default=30)
    ap.add_argument("--ntraj", type=int, default=2000)
    ap.add_argument("--traj-batch", type=int, default=250)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--heldout", type=int, default=64)
    ap.add_argument("--base-order", type=int, default=1, choices=[1, 2, 3],
                    help="learned workload: the extrapolation order the trained network corrects")
    ap.add_argument("--weights", default="", help="learned: load trained weights instead of training")
    ap.add_argument("--save-weights", default="")
    ap.add_argument("--lin-rtol", type=float, default=1e-10, help="newton: PCG relative tolerance")
    ap.add_argument("--femsolver", action="store_true", help="newton: also run FEMSolver (reference semantics)")
    ap.add_argument("--linear", default="mg", choices=["mg", "bj"],
                    help="newton: linear solver of the Newton steps (p-multigrid or block-Jacobi CG)")
    ap.add_argument("--copy-mib", type=int, default=0, help="assembly, newton: run one D2D copy of this size first "
                                                             "(calibrates PMC byte counters)")
    ap.add_argument("--modes", default="fixed,adaptive,fast",
                    help="newton: which solves of the twisted cube to run (fixed, adaptive, fast)")
    ap.add_argument("--newton-members", type=int, default=16,
                    help="assembly: held-out members of the Newton comparison in the default line (0 = skip)")
    ap.add_argument("--newton-n", type=int, default=24,
                    help="assembly: cells per edge of the twisted hex p=4 cube's Newton solve in the default line "
                         "(0 = skip)")
    ap.add_argument("--numbering", default="first", choices=["first", "reference"],
                    help="assembly: global node numbering of the p=4 mesh (first appearance, or the reference's "
                         "GetHighOrderMesh numbering, built on the device)")
    ap.add_argument("--nmesh", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16"])
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the assembly would shard by "
              "node-row ownership with halo elements). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "newton":
        return bench_newton(args)

    import torch
    from florence_amd import _lib
    from florence_amd.assembly import DeviceMeshData, assemble_device, residual_device
    from florence_amd.workloads import cube_hex_p4, twisted_state

    if args.stiffness_minw:
        _lib.check(_lib.load().fl_set_tuning(_lib.FL_TUNE_STIFF_MINW, args.stiffness_minw), "fl_set_tuning")
    if args.dense:
        _lib.check(_lib.load().fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 1), "fl_set_tuning")
    elif args.sf_form:
        _lib.check(_lib.load().fl_set_tuning(_lib.FL_TUNE_STIFF_SF, args.sf_form), "fl_set_tuning")
    if args.gather_form:
        _lib.check(_lib.load().fl_set_tuning(_lib.FL_TUNE_GATHER, args.gather_form), "fl_set_tuning")
    if args.stiffness_variant >= 0:
        _lib.check(_lib.load().fl_set_tuning(_lib.FL_TUNE_STIFF_VARIANT_BASE + 4, args.stiffness_variant + 1),
                   "fl_set_tuning")
    t_setup = time.perf_counter()
    mesh, fs, mat = cube_hex_p4(args.n, numbering=args.numbering)
    x_host = twisted_state(mesh, 15.0)
    dm = DeviceMeshData(mesh, fs, build_indices=False)
    x = torch.as_tensor(x_host, device="cuda")
    V = torch.empty(dm.nnz, dtype=torch.float64, device="cuda")
    T = torch.empty(3 * dm.nnode, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    print("setup {:.1f}s nelem={} nnode={} nnz={} max_degree={}".format(
        t_setup, dm.nelem, dm.nnode, dm.nnz, dm.max_degree), file=sys.stderr, flush=True)

    if args.copy_mib:
        copy_calibration(torch, args.copy_mib)
    direct = not args.no_direct
    for _ in range(args.warmup):
        assemble_device(dm, mat, x, V, T, direct=direct)
    torch.cuda.synchronize()
    status = dm.status.cpu().numpy().tolist()

    _lib.profile_enable(True)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    with no_gc():
        start.record()
        for _ in range(args.steps):
            assemble_device(dm, mat, x, V, T, direct=direct)
        stop.record()
        torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)
    calls = _lib.profile_read()
    _lib.profile_enable(False)
    phases = np.array([c[1] for c in calls if c[0] == 1])
    phase_ms = phases.mean(axis=0) if phases.size else np.full(4, np.nan)
    ms_per_step = total_ms / args.steps
    value = dm.nelem / (ms_per_step / 1e3)

    sf = dm.tensor_basis is not None and not args.dense
    direct = direct and sf and dm.scatter is not None and args.sf_form not in (3, 5)
    p1, q1 = int(round(dm.npe ** (1 / 3.))), int(round(dm.ngauss ** (1 / 3.)))
    flops_exec = (sf_flops_per_element(p1, q1) if sf else stiffness_flops_per_element(dm.npe, dm.ngauss)) * dm.nelem
    achieved = flops_exec / (phase_ms[1] / 1e3) / 1e12
    dense_b = dense_b_flops_per_element(dm.npe, dm.ngauss) * dm.nelem
    # the dominant kernel against the assembly's compulsory bytes: the dense-B flop
    # count cannot bound a sum-factorised kernel (it executes 37x fewer flops), so the
    # fraction is taken on bytes and the executed-flop rate is kept beside it
    cbytes = compulsory_bytes(dm.nelem, dm.npe, dm.nnode, dm.nnz)
    cgbs = cbytes / (phase_ms[1] / 1e3) / 1e9
    stiff_roof = {"bound": "hbm", "achieved": round(cgbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(cgbs / HBM_PEAK_GBS, 4),
                  "traffic": measured_traffic(PMC_SUMMARY, "stiffness"), "traffic_source": PMC_SOURCE,
                  "algorithmic_bytes_per_assembly": cbytes,
                  "kernel": (DIRECT_KERNEL if direct else SF_KERNELS[args.sf_form] if sf else
                             "stiffness_kernel<128,4,4,2,2> (off-diagonal + triangular diagonal launches)"),
                  "kernel_ms": round(float(phase_ms[1]), 4),
                  "mfma_view": {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                                "flops_per_assembly": flops_exec,
                                "flops": "executed (sum-factorised count)" if sf else "executed (dense blocks)"},
                  "dense_b_equivalent_tflops": round(dense_b / (phase_ms[1] / 1e3) / 1e12, 3),
                  "dense_block_equivalent_tflops": round(stiffness_flops_per_element(dm.npe, dm.ngauss) * dm.nelem
                                                         / (phase_ms[1] / 1e3) / 1e12, 3)}
    if direct:
        # one launch per colour: "per launch" figures are per assembly (all colours)
        stiff_roof["launches_per_assembly"] = dm.ncolour
        sbytes = direct_scatter_bytes(dm.nelem, dm.npe, dm.ngauss, dm.nnz, dm.scatter.numel() * 2)
        sgbs = sbytes / (phase_ms[1] / 1e3) / 1e9
        stiff_roof["scatter_view"] = {"bound": "hbm", "achieved": round(sgbs, 1), "peak": HBM_PEAK_GBS,
                                      "unit": "GB/s", "frac": round(sgbs / HBM_PEAK_GBS, 4),
                                      "bytes_per_assembly": sbytes,
                                      "bytes": "C fields, scatter table, V written, add-loads of later colours"}
    # the default factorised kernel writes element-interior rows itself (nodes of one
    # element, the lowest gather class when its degree equals npe)
    fused = sf and not direct and args.sf_form not in (3, 5) and dm.bucket_degree[0] == dm.npe
    n_interior = int(((dm.node_elem_ptr[1:] - dm.node_elem_ptr[:-1]) == 1).sum().item()) if fused else 0
    if direct:
        tbytes = 8 * (dm.nelem * 3 * dm.npe + 3 * dm.nnode)
        tgbs = tbytes / (phase_ms[3] / 1e3) / 1e9
        gather_roof = {"bound": "hbm", "achieved": round(tgbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(tgbs / HBM_PEAK_GBS, 4), "bytes_per_launch": tbytes,
                       "kernel": "gather_traction_kernel (T only: K is complete after the stiffness launches)",
                       "kernel_ms": round(float(phase_ms[3]), 4)}
    else:
        gbytes = gather_bytes(dm.nelem, dm.npe, dm.nnode, dm.nnz, n_interior)
        gather_gbs = gbytes / (phase_ms[3] / 1e3) / 1e9
        gather_roof = {"bound": "hbm", "achieved": round(gather_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(gather_gbs / HBM_PEAK_GBS, 4),
                       "traffic": measured_traffic("assembly_hex_p4_24cube_pmc.md", "gather"),
                       "kernel": "gather_kernel (all degree classes of one assembly)", "bytes_per_launch": gbytes,
                       "interior_nodes_written_by_stiffness": n_interior,
                       "kernel_ms": round(float(phase_ms[3]), 4)}
    # the other data path on the same mesh for comparison: element blocks stored, then
    # gathered per node (fl_df_assemble_mt)
    other_path = None
    if direct:
        assemble_device(dm, mat, x, V, T, direct=False)
        torch.cuda.synchronize()
        o0, o1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with no_gc():
            o0.record()
            for _ in range(3):
                assemble_device(dm, mat, x, V, T, direct=False)
            o1.record()
            torch.cuda.synchronize()
        other_path = {"form": "element blocks stored, then gathered per node", "ms_per_step": round(
            o0.elapsed_time(o1) / 3, 4)}
    # the precision form's tangents on the same mesh: FP32 (sum-factorised, FP32 MFMA and
    # accumulation) and BF16 (dense BF16 MFMA) element contractions, blocks gathered into
    # the FP64 CSR; kinematics, stress and T in FP64
    precision_forms = {}
    V32 = torch.empty(dm.nnz, dtype=torch.float32, device="cuda")
    for prec, csr in (("fp32", "fp64"), ("fp32", "fp32"), ("bf16", "fp64")):
        Vp = V32 if csr == "fp32" else V
        assemble_device(dm, mat, x, Vp, T, precision=prec, csr=csr)
        torch.cuda.synchronize()
        _lib.profile_enable(True)
        q0, q1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with no_gc():
            q0.record()
            for _ in range(3):
                assemble_device(dm, mat, x, Vp, T, precision=prec, csr=csr)
            q1.record()
            torch.cuda.synchronize()
        pcalls = _lib.profile_read()
        _lib.profile_enable(False)
        pph = np.array([c[1] for c in pcalls if c[0] == 1]).mean(axis=0)
        key = prec + ("+csr32" if csr == "fp32" else "")
        precision_forms[key] = {"ms_per_step": round(q0.elapsed_time(q1) / 3, 4),
                                 "elements_per_s": round(dm.nelem / (q0.elapsed_time(q1) / 3e3), 1),
                                 "phase_ms": {k: round(float(v), 4) for k, v in
                                              zip(("qp", "stiffness", "traction", "gather"), pph)}}
        print("{} tangent: {:.3f} ms per assembly".format(key, precision_forms[key]["ms_per_step"]),
              file=sys.stderr, flush=True)
    del V32
    # the practical read+write rate of this GPU: a 4 GiB device-to-device copy, same
    # clock (HIP events), for comparison with the HBM-bound phases
    del V
    copy_gbs = device_copy_gbs(torch, 4 << 30)
    gather_roof["copy_gbs_measured"] = round(copy_gbs, 1)
    if not direct:
        gather_roof["frac_of_copy"] = round(gather_gbs / copy_gbs, 4)
    dominant_is_stiffness = phase_ms[1] >= phase_ms[3]
    print("phases ms: qp {:.3f} stiffness {:.3f} traction {:.3f} gather {:.3f}".format(*phase_ms),
          file=sys.stderr, flush=True)

    # matrix-free residual-only pass on the same mesh (explicit DF internal force)
    dm.release_workspace()
    Tres = torch.empty(3 * dm.nnode, dtype=torch.float64, device="cuda")
    residual_device(dm, mat, x, Tres)
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    r0, r1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with no_gc():
        r0.record()
        for _ in range(args.steps):
            residual_device(dm, mat, x, Tres)
        r1.record()
        torch.cuda.synchronize()
    res_ms = r0.elapsed_time(r1) / args.steps
    rcalls = _lib.profile_read()
    _lib.profile_enable(False)
    rph = np.array([c[1][:3] for c in rcalls if c[0] == 2]).mean(axis=0)
    # residual-only algorithmic work: G * (108 npe + n + 300) flops per element (dense-B count)
    res_flops = dm.ngauss * (108 * dm.npe + 3 * dm.npe + 300) * dm.nelem
    residual_only = {"elements_per_s": round(dm.nelem / (res_ms / 1e3), 1), "ms_per_pass": round(res_ms, 4),
                     "phase_ms": {"qp": round(float(rph[0]), 4), "traction": round(float(rph[1]), 4),
                                  "gather": round(float(rph[2]), 4)},
                     "dense_b_equivalent_tflops": round(res_flops / (res_ms / 1e3) / 1e12, 3),
                     "max_abs_diff_vs_full_T": float((Tres - T).abs().max())}

    cpu = None
    if args.cpu_sample > 0:
        cpu = cpu_baseline(mesh, fs, mat, x_host, args.cpu_sample)
    sizes = {"nelem": dm.nelem, "nnode": dm.nnode, "nnz": dm.nnz, "ngauss": dm.ngauss}
    ncolour = dm.ncolour
    del dm, x, T, Tres
    torch.cuda.empty_cache()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    newton = newton_comparison(args, log) if args.newton_members else None
    newton_c2 = None
    if args.newton_n > 0:
        modes = twisted_cube_newton(args.newton_n, "mg", 1e-10, log)
        plain_fixed = modes["fp64+previous+fixed"]["wall_s"]
        plain_adapt = modes["fp64+previous+adaptive"]["wall_s"]
        fast = modes["fp32+poly+adaptive+csr32"]["wall_s"]
        newton_c2 = {"workload": "unit cube {0}^3 hex p=4 (jitter 0.03h), 6^3 GLL, MR(0.5,0.5,20), x=0 clamped, x=1 "
                                 "twisted 30 deg, 10 increments, |R_free| <= 1e-8 R0, p-multigrid flexible CG "
                                 "(fixed: linear rtol 1e-10; adaptive: forcing terms)".format(args.newton_n),
                     "mixed": "FP32 tangent, FP32-stored CSR, polynomial-extrapolation start, adaptive",
                     "plain_fp64_fixed_s": plain_fixed, "plain_fp64_adaptive_s": plain_adapt, "mixed_fp32_s": fast,
                     "speedup_equal_policy": round(plain_adapt / fast, 3),
                     "speedup_vs_plain_fixed": round(plain_fixed / fast, 3), "modes": modes}

    out = {
        "metric": "K+residual assembly elements/s & Newton time-to-1e-8 vs plain FP64 on GPU",
        "value_is": "K+residual assembly elements/s (DF, hex p=4, FP64); the Newton half is newton_twisted_cube",
        "value": round(value, 1),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "unit cube 24^3 hex p=4 (13824 elements, 125 nodes/element), 6^3 Gauss-Lobatto "
                               "points, Mooney-Rivlin(0.5,0.5,20), twisted state, K+T into reference CSR "
                               "(nnz {})".format(sizes["nnz"]), "numbering": args.numbering, **sizes},
        "roofline": stiff_roof if dominant_is_stiffness else gather_roof,
        "roofline_other_phase": gather_roof if dominant_is_stiffness else stiff_roof,
        "phase_ms": {"qp": round(float(phase_ms[0]), 4), "stiffness": round(float(phase_ms[1]), 4),
                     "traction": round(float(phase_ms[2]), 4), "gather": round(float(phase_ms[3]), 4)},
        "cpu_baseline": cpu,
        "scatter": "direct (colour-ordered, {} colours)".format(ncolour) if direct else "element blocks + gather",
        "other_data_path": other_path,
        "precision_forms": precision_forms,
        "residual_only": residual_only,
        "newton": newton,
        "newton_twisted_cube": newton_c2,
        "status": status,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
