"""Benchmark of the SPH step on one MI355X (libmfl.so via the oclFluid3D program object).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1m|small] [--no-cpu]

Workload: the 1,048,576-particle dam break (128x64x128 jittered lattice at 0.5h, h = 0.02, box
[0,4]^3, fp32), inputs resident in HBM. A step is one oclFluid3D::compute(): hash + bucket clear,
sort, reorder + bounds, density, forces + clip box + integration. Prints one JSON line; its
secondary legs are the batched scenes, the adjoint, the particle-vs-mesh collision (BVH build +
clCollideBVH on the same 1M particles), the 2D sibling solver oclFluid3Dnext at 1M particles, the
standalone key/value sort of 16M cell keys and the 16k dam break (eager against hipGraph replay).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
# VALU issue: 256 CUs x 4 SIMDs at 2.4 GHz. A wave64 instruction holds its SIMD for 4 cycles when it is an
# FP32 add, mul or fma (or a 32-bit shift) and for 2 when it is a 32-bit integer add, bitwise, compare or
# select (measured on MI355X with independent chains of one instruction: 603-608 against 1145-1195 G/s).
VALU_SIMDS, VALU_GHZ = 256 * 4, 2.4
VALU_PEAK_G = VALU_SIMDS * VALU_GHZ / 4  # G wave64 instructions/s, every instruction at 4 cycles

# Algorithmic HBM bytes per particle of each stage: the reference dataflow with the minimum
# materialisation its API exposes (sorted buffers, index and state), each array touched once;
# neighbour re-reads are cache traffic and not counted. 240 B per particle in all, plus 4 MiB per step
# (the 65,536-entry clear and one read of both 2^18 bucket tables).
STAGE_BYTES_PER_PARTICLE = {
    "hash": 16 + 4 + 4,                 # R pos, W key, W index
    "sort": 3 * (8 + 8) + 4,            # 24-bit keys, 3 passes of 8 bits, one histogram read
    "reorder": 4 + 32 + 32,             # R index, R pos/vel, W sorted pos/vel
    "bounds": 4,                        # R key
    "density": 16 + 8,                  # R sorted pos, W density/pressure
    "forces": 32 + 4 + 32,              # R sorted pos/vel, R index, W pos/vel
}
STEP_BYTES_PER_LAUNCH = 4 << 20
STAGE_BYTES_PER_LAUNCH = {
    "density": 2 * 262144 * 4,          # one read of both bucket tables
    "forces": 2 * 262144 * 4,
}
TUNING_KEYS = ("neighbour_pass", "density_lanes", "tile_depth", "block_waves", "sort_items", "force_chunks")


def _profiled(name: str, workload: str, tuning: dict):
    """A committed counter summary's table for a workload, when it was measured at exactly the launch
    shape `tuning` (all six tuning fields and the traversal order)."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as fh:
            table = json.load(fh)
    except (OSError, ValueError):
        return None
    shape = table.get("shape", {})
    if any(shape.get(k) != tuning.get(k) for k in TUNING_KEYS + ("traversal",)):
        return None
    return table.get(workload)


def measured_traffic(stage: str, workload: str, tuning: dict):
    """HBM bytes per launch of a stage's kernel from the committed rocprofv3 counter summary
    (profiles/traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy calibration);
    None when that kernel and workload were not measured."""
    t = _profiled("traffic.json", workload, tuning)
    return t[stage]["bytes_per_launch"] if t and stage in t else None


def measured_valu(stage: str, workload: str, tuning: dict):
    """(VALU wave64 instructions per launch, FP32 share of them) of a stage's kernel from the committed
    rocprofv3 SQ counter summary (profiles/valu.json: SQ_INSTS_VALU per dispatch and the FP32 classes'
    share, SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F32); None when that kernel and workload were not measured at
    this launch shape (the share is None when the class counters were not taken)."""
    t = _profiled("valu.json", workload, tuning)
    if not t or stage not in t:
        return None
    return t[stage]["valu_per_wave"] * t[stage]["waves"], t[stage].get("fp32_share")


def progress(msg: str) -> None:
    print(f"bench: {msg}", file=sys.stderr, flush=True)


_HIP = None


def _copy_d2d(dst: int, src: int, nbytes: int, stream) -> None:
    """hipMemcpyAsync device to device on `stream`, through the HIP runtime torch has loaded."""
    global _HIP
    if _HIP is None:
        import ctypes

        import torch
        lib = os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so")
        _HIP = ctypes.CDLL(lib if os.path.exists(lib) else "libamdhip64.so")
        _HIP.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                        ctypes.c_void_p]
        _HIP.hipMemcpyAsync.restype = ctypes.c_int
    rc = _HIP.hipMemcpyAsync(dst, src, nbytes, 3, stream.cuda_stream)  # 3 = hipMemcpyDeviceToDevice
    if rc != 0:
        raise RuntimeError(f"hipMemcpyAsync failed with error {rc}")


class DeviceSnapshot:
    """Device buffers ((pointer, bytes) pairs) copied device to device, and put back the same way:
    restoring them takes microseconds and no host round trip, so the card stays at the clocks of the
    work before it. `after_restore` runs after each restore."""

    def __init__(self, buffers, stream, after_restore=None):
        import torch
        self.bufs, self.stream, self.after = list(buffers), stream, after_restore
        self.copies = [torch.empty(nbytes, dtype=torch.uint8, device="cuda") for _, nbytes in self.bufs]
        for (ptr, nbytes), c in zip(self.bufs, self.copies):
            _copy_d2d(c.data_ptr(), ptr, nbytes, stream)
        stream.synchronize()

    def restore(self) -> None:
        for (ptr, nbytes), c in zip(self.bufs, self.copies):
            _copy_d2d(ptr, c.data_ptr(), nbytes, self.stream)
        if self.after:
            self.after()
        self.stream.synchronize()


def program_snapshot(f, stream) -> DeviceSnapshot:
    """A single-scene program's state: positions, velocities, forces and both bucket tables. A restore
    also makes the next step rebuild the Morton order, so every leg started from the snapshot refreshes
    it at the same steps."""
    bufs = [(b.ptr, b.nbytes) for b in (f.bfPosition, f.bfVelocity, f.bfForce, f.bfCellStart, f.bfCellEnd)]
    return DeviceSnapshot(bufs, stream, lambda: f.setTraversal(*f.getTraversal()))


def hold_clocks(step, stream, reset=None, seconds: float = 0.3) -> int:
    """Calls step() untimed for `seconds`, in rounds of 16 calls that each start with reset() (restoring
    the state the timed steps start from, so the untimed steps cost what the timed ones do), and returns
    the call count. The card lowers its clocks within milliseconds of idling and needs tens of
    milliseconds of continuous work to raise them again: 20 steps of the 1M workload take 0.22-0.24 ms
    each right after a host round trip or an idle pause, 0.21 ms after this."""
    t0, k = time.perf_counter(), 0
    while time.perf_counter() - t0 < seconds:
        if reset:
            reset()
        for _ in range(16):
            step()
        stream.synchronize()
        k += 16
    return k


def host_cores() -> tuple:
    """(cores this process may run on, cores of the host)."""
    try:
        mine = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        mine = os.cpu_count() or 1
    return mine, os.cpu_count() or mine


def cpu_baseline(sc, prm, steps: int, steps_one_core: int) -> dict:
    """The oracle (a plain-C OpenMP restatement of the reference kernels) on every core this process may
    run on, then on one core, each over a bounded number of steps from the workload's initial state."""
    from oracle import oracle as O
    O.build()
    mine, total = host_cores()
    op = O.Params(*prm.values())

    def run(threads: int, k: int):
        used = O.set_threads(threads)
        fo = O.Fluid(sc.pos, sc.vel, sc.force, op, box_min=sc.box_min, box_max=sc.box_max)
        fo.step()  # warm-up
        t0 = time.perf_counter()
        fo.steps(k)
        return used, time.perf_counter() - t0

    used, dt = run(mine, steps)
    _, dt1 = run(1, steps_one_core)
    return {"value": sc.n * steps / dt, "unit": "particle-steps/s", "cores": used, "kind": "port",
            "host_cores": total, "one_core_value": sc.n * steps_one_core / dt1,
            "sample": f"{steps} steps of the same {sc.n}-particle workload after 1 warm-up step, "
                      f"oracle/sph_oracle.c with {used} OpenMP threads ({dt:.1f} s; the host has {total} cores, "
                      f"{mine} of them available to this process); one_core_value: {steps_one_core} steps on "
                      f"1 thread ({dt1:.1f} s)"}


def batched_leg(steps: int, seq_steps: int, num_scenes: int = 256) -> dict:
    """256 independent scenes (log-uniform 2,048..65,536 particles) advanced together (launch shape
    autotuned, hipGraph replay), against the same scenes stepped one after another by single-scene
    programs with the same launch shape on the same stream."""
    import numpy as np
    import torch

    import libcl_amd as L
    from libcl_amd import scenes

    scs = scenes.batched_scenes(num=num_scenes, seed=5, steps=steps)
    prm = L.params_from(scs[0].params)
    total = sum(sc.n for sc in scs)
    stream = torch.cuda.Stream()
    b = L.FluidBatch([sc.n for sc in scs])
    assert b.setParameters(prm) == 1 and b.setBox((0, 0, 0, 0), (1, 1, 1, 0)) == 1
    assert b.upload(np.concatenate([sc.pos for sc in scs]), np.concatenate([sc.vel for sc in scs]),
                    np.concatenate([sc.force for sc in scs])) == 1
    res = b.autotune(steps=3, stream=stream)
    assert res is not None, L.last_error()
    b.setUseGraph(True)
    for _ in range(3):
        assert b.compute(stream) == 1, L.last_error()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.compute(stream)
    torch.cuda.synchronize()
    batched_ms = (time.perf_counter() - t0) * 1e3 / steps

    progs = []
    for sc in scs:
        f = L.oclFluid3D()
        assert f.compile() == 1 and f.setParticleCount(sc.n) == 1
        assert f.setTuning(**res[0]) == 1, L.last_error()  # the batch's launch shape
        f.setParameters(prm)
        f.setBox(sc.box_min, sc.box_max)
        assert f.upload(sc.pos, sc.vel, sc.force) == 1
        progs.append(f)
    out = {}
    for mode in ("eager", "graph"):
        for f in progs:
            f.setUseGraph(mode == "graph")
            f.compute(stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(seq_steps):
            for f in progs:
                f.compute(stream)
        torch.cuda.synchronize()
        out[mode] = (time.perf_counter() - t0) * 1e3 / seq_steps
    del progs
    return {"scenes": num_scenes, "particles": total, "steps": steps, "tuning": res[0], "ms_per_step": batched_ms,
            "particle_steps_per_s": total / (batched_ms / 1e3),
            "sequential_ms_per_step": out["eager"], "sequential_graph_ms_per_step": out["graph"],
            "sequential_steps": seq_steps, "gain": out["eager"] / batched_ms,
            "gain_vs_graphed_sequential": out["graph"] / batched_ms}


def adjoint_leg(steps: int = 200, every: int = 20) -> dict:
    """Differentiable case: 65,536 particles (64x32x32), 200 steps, checkpoints every 20. Times the
    checkpointed forward rollout and the backward pass (segment recompute + reverse steps) on one
    stream; the loss is mean |x_T - target|^2 in FP64."""
    import numpy as np
    import torch

    import libcl_amd as L
    from libcl_amd import scenes

    sc, target = scenes.config_grad()
    prm = L.params_from(sc.params)
    f = L.oclFluid3D()
    assert f.compile() == 1 and f.setParticleCount(sc.n) == 1
    f.setParameters(prm)
    f.setBox(sc.box_min, sc.box_max)
    adj = L.Adjoint(f, steps, every, 192)
    stream = torch.cuda.Stream()
    tgt = torch.from_numpy(target).cuda()

    def run():
        assert f.upload(sc.pos, sc.vel, sc.force) == 1
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        assert adj.forward(stream) == 1, L.last_error()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        xt = torch.from_numpy(f.state()["pos"]).cuda()
        xd = xt[:, :3].double()
        loss = ((xd - tgt) ** 2).sum(1).mean()
        gbar = torch.zeros_like(xt)
        gbar[:, :3] = (2.0 * (xd - tgt) / sc.n).float()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        assert adj.backward(gbar.contiguous(), None, stream) == 1, L.last_error()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        return t1 - t0, t3 - t2, loss.item()

    run()  # warm-up (allocations, first launches)
    fwd, bwd, loss = run()
    pg = adj.param_grads()
    # plain steps of the same program without checkpoints, for reference
    assert f.upload(sc.pos, sc.vel, sc.force) == 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        f.compute(stream)
    torch.cuda.synchronize()
    plain = time.perf_counter() - t0
    return {"particles": sc.n, "steps": steps, "checkpoint_every": every, "loss": loss,
            "forward_ms_per_step": fwd * 1e3 / steps, "backward_ms_per_step": bwd * 1e3 / steps,
            "plain_ms_per_step": plain * 1e3 / steps, "adjoint_over_forward": bwd / fwd,
            "grad_viscosity": pg["viscosity"], "grad_stiffness": pg["stiffness"], "grad_density": pg["density"]}


def collide_leg(launches: int = 20, cpu: bool = True) -> dict:
    """Particle-vs-mesh collision (oclBvhTrimesh + clCollideBVH) on the 1,048,576-particle block falling
    onto a 9,216-triangle floor (collide_1m_floor9k), with clInitFluid's Params: BVH build time (the
    build ends by reading the root back, so it is timed by the host clock), the build of a
    1,048,576-triangle floor, and clCollideBVH over all particles, each launch from the same input
    (restored between launches, outside the HIP events that time the kernel)."""
    import numpy as np
    import torch

    import libcl_amd as L
    from libcl_amd import scenes

    sc, mesh = scenes.config_collide_1m()
    prm = L.default_params()
    b = L.oclBvhTrimesh()
    assert b.compile() == 1, L.last_error()
    stream = torch.cuda.Stream()

    def build_ms(m, reps):
        vtx = torch.from_numpy(m.vertex).cuda()
        idx = torch.from_numpy(m.index.view(np.int32)).cuda()
        for _ in range(2):
            assert b.compute(vtx, idx, stream) == 1, L.last_error()
        t0 = time.perf_counter()
        for _ in range(reps):
            assert b.compute(vtx, idx, stream) == 1, L.last_error()
        return (time.perf_counter() - t0) * 1e3 / reps, vtx, idx

    big = scenes.floor_mesh(1024, 512, (0.0, 0.0), (4.0, 2.0), 0.3, 0.001, 7)
    big_ms, _, _ = build_ms(big, 5)
    ms, vtx, idx = build_ms(mesh, 20)
    P0, V0 = torch.from_numpy(sc.pos).cuda(), torch.from_numpy(sc.vel).cuda()
    F0 = torch.from_numpy(sc.force).cuda()
    P, V, F = P0.clone(), V0.clone(), F0.clone()
    times = []
    for k in range(launches + 2):
        with torch.cuda.stream(stream):
            P.copy_(P0)
            V.copy_(V0)
            F.copy_(F0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        assert L.collide_bvh(F, P, V, prm, b, idx, vtx, stream) == 1, L.last_error()
        e1.record(stream)
        times.append((e0, e1))
    torch.cuda.synchronize()
    us = [a.elapsed_time(z) * 1e3 for a, z in times[2:]]
    avg = sum(us) / len(us)
    contacts = int((V[:, :3] == 0).all(1).sum().item())
    out = {"workload": sc.name, "particles": sc.n, "triangles": mesh.triangles, "build_ms": ms,
           "build_1m_triangles_ms": big_ms, "collide_us_per_launch": avg, "particles_per_s": sc.n / (avg * 1e-6),
           "contacts": contacts, "launches": launches}
    if cpu:
        from oracle import oracle as O
        O.build()
        threads = O.set_threads(host_cores()[0])
        nodes = b.nodes()
        pos, vel, force = sc.pos.copy(), sc.vel.copy(), sc.force.copy()
        op = O.Params(*prm.values())
        O.collide_bvh(pos[:1024], vel[:1024], force[:1024], op, nodes, b.getRootNode(), mesh.index, mesh.vertex)
        t0 = time.perf_counter()
        O.collide_bvh(pos, vel, force, op, nodes, b.getRootNode(), mesh.index, mesh.vertex)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": pos.shape[0] / dt, "unit": "particles/s", "cores": threads, "kind": "port",
                               "sample": f"one clCollideBVH pass over all {pos.shape[0]} particles, "
                                         f"oracle/bvh_oracle.c with {threads} OpenMP threads ({dt:.2f} s)"}
    return out


def next_leg(steps: int = 50, cpu: bool = True) -> dict:
    """oclFluid3Dnext (the 2D dense-grid sibling solver) on 1,048,576 particles (config_next_1m): steps
    timed with HIP events on one stream, inputs resident; the oracle times one step on the host."""
    import torch

    import libcl_amd as L
    from libcl_amd import scenes

    st = scenes.config_next_1m()
    n = st.shape[0]
    f = L.oclFluid3Dnext()
    assert f.compile() == 1 and f.setParticleCount(n) == 1, L.last_error()
    assert f.upload(st) == 1
    stream = torch.cuda.Stream()
    for _ in range(3):
        assert f.compute(stream) == 1, L.last_error()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        f.compute(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    out = {"workload": "next_1m_block", "particles": n, "steps": steps, "ms_per_step": ms,
           "particle_steps_per_s": n / (ms * 1e-3)}
    if cpu:
        from oracle import oracle as O
        O.build()
        threads = O.set_threads(host_cores()[0])
        fo = O.Next(st, O.next_default_params())
        fo.step()
        t0 = time.perf_counter()
        fo.step()
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
                               "sample": f"one step of the same {n}-particle state, oracle/next_oracle.c with "
                                         f"{threads} OpenMP threads in its neighbour passes ({dt:.2f} s)"}
    return out


def small_leg(steps: int = 1000, reps: int = 3, cpu: bool = True) -> dict:
    """The 16,384-particle dam break in the unit box (sph_16k_unit_box), its 1000 steps from the initial
    state with the autotuned launch shape, as eager launches and as hipGraph replays on one stream,
    alternating, `reps` times each (the best run of each counts), each run after 0.3 s of untimed
    stepping (hold_clocks): at this size a step is launch-bound, so the figure is particle-steps/s and
    the graph gain."""
    import numpy as np
    import torch

    import libcl_amd as L
    from libcl_amd import scenes

    sc = scenes.config_small()
    prm = L.params_from(sc.params)
    stream = torch.cuda.Stream()
    progs = {}
    tuning = None
    for mode in ("eager", "graph"):
        f = L.oclFluid3D()
        assert f.compile() == 1 and f.setParticleCount(sc.n) == 1, L.last_error()
        f.setParameters(prm)
        f.setBox(sc.box_min, sc.box_max)
        if tuning is None:  # the launch shape the autotuner picks from the initial state, for both modes
            assert f.upload(sc.pos, sc.vel, sc.force) == 1
            res = f.autotune(steps=3, stream=stream)
            assert res is not None, L.last_error()
            tuning = res[0]
        else:
            assert f.setTuning(**tuning) == 1, L.last_error()
        assert f.setUseGraph(mode == "graph") == 1
        progs[mode] = f
    empty = np.full(262144, 0xFFFFFFFF, np.uint32)
    snaps = {}
    for mode, f in progs.items():  # the initial state, restored device to device before each timed run
        assert f.upload(sc.pos, sc.vel, sc.force) == 1
        assert f.bfCellStart.write(empty) == 1 and f.bfCellEnd.write(empty) == 1
        torch.cuda.synchronize()
        snaps[mode] = program_snapshot(f, stream)
    best = {"eager": float("inf"), "graph": float("inf")}
    for _ in range(reps):
        for mode, f in progs.items():
            hold_clocks(lambda f=f: f.compute(stream), stream, snaps[mode].restore)
            snaps[mode].restore()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                f.compute(stream)
            torch.cuda.synchronize()
            best[mode] = min(best[mode], (time.perf_counter() - t0) * 1e3 / steps)
    out = {"workload": sc.name, "particles": sc.n, "steps": steps, "reps": reps, "tuning": tuning}
    for mode, ms in best.items():
        out[f"{mode}_ms_per_step"] = ms
        out[f"{mode}_particle_steps_per_s"] = sc.n / (ms * 1e-3)
    out["graph_gain"] = best["eager"] / best["graph"]
    if cpu:
        from oracle import oracle as O
        O.build()
        threads = O.set_threads(host_cores()[0])
        op = O.Params(*prm.values())
        fo = O.Fluid(sc.pos, sc.vel, sc.force, op, box_min=sc.box_min, box_max=sc.box_max)
        t0 = time.perf_counter()
        fo.steps(steps)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": sc.n * steps / dt, "unit": "particle-steps/s", "cores": threads,
                               "kind": "port", "sample": f"the same {steps} steps, oracle/sph_oracle.c with "
                                                         f"{threads} OpenMP threads ({dt:.1f} s)"}
    return out


SORT_BYTES_PER_PAIR = 52  # 3 passes x (R 8 + W 8) + one 4-byte histogram read per key


def sort_leg(reps: int = 10, cpu: bool = True) -> dict:
    """oclRadixSort on 16,777,216 (key, value) pairs by bits [0, 24): keys over 2^24 cells, uniform
    (seed 3) and skewed (half the keys in a hot 10% of the cells, chosen with seed 4), values 0..N-1.
    Every timed sort starts from the same unsorted input, restored between sorts outside the HIP events
    that time the sort, after 0.3 s of untimed sorts (hold_clocks); the sorted output is checked for order
    and for the value permutation."""
    import numpy as np
    import torch

    import libcl_amd as L
    from libcl_amd import scenes

    n = 1 << 24
    s = L.oclRadixSort()
    assert s.compile() == 1, L.last_error()
    out = {"workload": "sort_16m_cells24", "pairs": n, "bits": [0, 24], "passes": 3,
           "alg_bytes_per_pair": SORT_BYTES_PER_PAIR, "reps": reps}
    keys_uniform = None
    for name, skew in (("uniform", False), ("skew", True)):
        keys = scenes.sort_keys(n, seed=3, skew=skew, bits=24)
        if not skew:
            keys_uniform = keys
        k0 = torch.from_numpy(keys.view(np.int32).copy()).cuda()
        v0 = torch.arange(n, dtype=torch.int32, device="cuda")
        k, v = k0.clone(), v0.clone()

        def one_sort():
            k.copy_(k0)
            v.copy_(v0)
            assert s.compute(k, v, 0, 24) == 1, L.last_error()

        hold_clocks(one_sort, torch.cuda.current_stream())
        times = []
        for _ in range(reps + 2):
            k.copy_(k0)
            v.copy_(v0)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            assert s.compute(k, v, 0, 24) == 1, L.last_error()
            e1.record()
            times.append((e0, e1))
        torch.cuda.synchronize()
        us = [a.elapsed_time(z) * 1e3 for a, z in times[2:]]
        avg = sum(us) / len(us)
        ku = k.to(torch.int64) & 0xFFFFFFFF  # the keys as unsigned values
        ordered = bool((ku[1:] >= ku[:-1]).all())
        perm = bool((torch.sort(v.to(torch.int64)).values == torch.arange(n, device="cuda")).all())
        gbs = SORT_BYTES_PER_PAIR * n / (avg * 1e-6) / 1e9
        out[name] = {"us_per_sort": avg, "pairs_per_s": n / (avg * 1e-6), "achieved_GBs": gbs,
                     "hbm_frac": gbs / HBM_PEAK_GBS, "ordered": ordered and perm}
    if cpu:
        from oracle import oracle as O
        O.build()
        ks, vs = keys_uniform.copy(), np.arange(n, dtype=np.uint32)
        t0 = time.perf_counter()
        O.std_stable_sort(ks, vs, 0, 24)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / dt, "unit": "pairs/s", "cores": 1, "kind": "port",
                               "sample": f"std::stable_sort of the {n} uniform pairs by bits [0, 24) "
                                         f"(oracle/std_stable_sort.cpp), one core ({dt:.2f} s)"}
        m = 4194300  # the reference's own sort takes fewer than 4,194,304 keys
        ks, vs = keys_uniform[:m].copy(), np.arange(m, dtype=np.uint32)
        t0 = time.perf_counter()
        assert O.radix_sort(ks, vs, 0, 24)
        dt = time.perf_counter() - t0
        out["cpu_reference_restated"] = {"value": m / dt, "unit": "pairs/s", "cores": 1, "kind": "port",
                                         "sample": f"the restated 4-bit reference sort (oracle/sph_oracle.c "
                                                   f"orc_radix_sort) on the first {m} uniform pairs ({dt:.2f} s)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="1m", choices=["1m", "small"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=80, help="oracle steps timed on all cores (~10 s)")
    ap.add_argument("--cpu-steps-one-core", type=int, default=6, help="oracle steps timed on one core (~10 s)")
    ap.add_argument("--order", default="morton", choices=["sorted", "caller", "morton"])
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-collide", action="store_true")
    ap.add_argument("--no-next", action="store_true")
    ap.add_argument("--no-sort", action="store_true")
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--only", default=None, choices=["batched", "adjoint", "collide", "next", "sort", "small"],
                    help="run one secondary leg alone and print its JSON (profiling)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB three times (counter calibration)")
    ap.add_argument("--pass", dest="npass", default="tile_kept", choices=["tile_kept", "tile_list", "tile", "list"],
                    help="neighbour passes: LDS-tile density + kept-entry forces, LDS-tile density + recorded-visit "
                         "forces, LDS tiles for both, or recorded visit lists for both")
    ap.add_argument("--tile-depth", type=int, default=None, help="default: the library's")
    ap.add_argument("--lanes", type=int, default=None, help="lanes per particle in the tile passes")
    ap.add_argument("--force-chunks", type=int, default=None, help="list chunks per trip in the force pass")
    ap.add_argument("--block-waves", type=int, default=None)
    ap.add_argument("--autotune", action="store_true",
                    help="time the candidate launch shapes first and keep the fastest (the timed shape then "
                         "may differ from the one the committed counter profiles were taken at)")
    ap.add_argument("--trajectory-steps", type=int, default=200,
                    help="also time the workload's whole trajectory from its initial state (0 = skip)")
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--sequential-steps", type=int, default=5)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the step is not sharded across GPUs",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import libcl_amd as L
    from libcl_amd import scenes

    if args.only == "batched":
        print(json.dumps(batched_leg(args.batched_steps, args.sequential_steps)), flush=True)
        return
    if args.only == "adjoint":
        print(json.dumps(adjoint_leg()), flush=True)
        return
    if args.only == "collide":
        print(json.dumps(collide_leg(cpu=not args.no_cpu)), flush=True)
        return
    if args.only == "next":
        print(json.dumps(next_leg(cpu=not args.no_cpu)), flush=True)
        return
    if args.only == "sort":
        print(json.dumps(sort_leg(cpu=not args.no_cpu)), flush=True)
        return
    if args.only == "small":
        print(json.dumps(small_leg(cpu=not args.no_cpu)), flush=True)
        return

    sc = scenes.config_1m() if args.config == "1m" else scenes.config_small()
    prm = L.params_from(sc.params)
    n = sc.n
    f = L.oclFluid3D()
    assert f.compile() == 1, L.last_error()
    assert f.setParticleCount(n) == 1, L.last_error()
    f.setParameters(prm)
    f.setBox(sc.box_min, sc.box_max)
    assert f.upload(sc.pos, sc.vel, sc.force) == 1
    assert f.setTraversal(args.order, 16) == 1
    shape = {"neighbour_pass": args.npass}
    for key, val in (("tile_depth", args.tile_depth), ("block_waves", args.block_waves),
                     ("density_lanes", args.lanes), ("force_chunks", args.force_chunks)):
        if val is not None:
            shape[key] = val
    assert f.setTuning(**shape) == 1, L.last_error()
    stream = torch.cuda.Stream()

    tuned = None
    if args.autotune:
        res = f.autotune(steps=3, stream=stream)
        assert res is not None, L.last_error()
        tuned = {"ms_per_step": res[1], "pick": res[0]}
    tuning = f.getTuning()  # with the traversal order (the autotuner may change it)

    # The card lowers its clocks within milliseconds of idling, so each timed leg below starts from a
    # device-side snapshot right after 0.3 s of untimed stepping (hold_clocks): the timed steps run as
    # in a simulation that steps continuously. `cold_ms_per_step` times the same steps after a 0.2 s pause.
    empty = np.full(262144, 0xFFFFFFFF, np.uint32)
    # one untimed step first (the program's first step sets up its buffers), then the initial state again
    assert f.compute(stream) == 1, L.last_error()
    assert f.upload(sc.pos, sc.vel, sc.force) == 1
    assert f.bfCellStart.write(empty) == 1 and f.bfCellEnd.write(empty) == 1
    torch.cuda.synchronize()
    initial = program_snapshot(f, stream)
    step = lambda: f.compute(stream)  # noqa: E731

    # the whole trajectory from the initial state: HIP events on the step stream at steps 0, 1, 10, 100
    # and the last (each step's cost grows as the flow develops)
    trajectory = None
    if args.trajectory_steps > 0:
        hold_clocks(step, stream, initial.restore)
        initial.restore()
        marks = sorted({0, 1, 10, 100, args.trajectory_steps} & set(range(args.trajectory_steps + 1)))
        events = {}
        for k in range(args.trajectory_steps + 1):
            if k in marks:
                events[k] = torch.cuda.Event(enable_timing=True)
                events[k].record(stream)
            if k < args.trajectory_steps:
                assert f.compute(stream) == 1, L.last_error()
        torch.cuda.synchronize()
        total_ms = events[marks[0]].elapsed_time(events[marks[-1]])
        trajectory = {"steps": args.trajectory_steps, "ms_per_step": total_ms / args.trajectory_steps,
                      "particle_steps_per_s": n * args.trajectory_steps / (total_ms / 1e3),
                      "segments_ms_per_step": {f"{a}-{b}": events[a].elapsed_time(events[b]) / (b - a)
                                               for a, b in zip(marks[:-1], marks[1:])}}

    initial.restore()
    for _ in range(args.warmup):
        assert f.compute(stream) == 1, L.last_error()
    stream.synchronize()
    start = program_snapshot(f, stream)  # every timed leg starts from this state

    # untimed stage profile: which kernel dominates the step
    f.setProfiling(L._native.STAGES)
    f.stageTime("hash", reset=True)
    for _ in range(5):
        f.compute(stream)
    stream.synchronize()
    stage_ms = {}
    for s in L._native.STAGES:
        ms, cnt = f.stageTime(s)
        stage_ms[s] = ms / max(cnt, 1)
    f.setProfiling([])
    f.stageTime("hash", reset=True)
    # the dominant per-step kernel ("order", the traversal refresh, runs every 16 steps, not per step)
    dominant = max((st for st in stage_ms if STAGE_BYTES_PER_PARTICLE.get(st)), key=stage_ms.get)

    def run_steps() -> float:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            f.compute(stream)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # timed region: K steps, HIP events around the dominant kernel on its stream
    hold_clocks(step, stream, start.restore)
    start.restore()
    f.setProfiling([dominant])
    f.stageTime(dominant, reset=True)
    elapsed = run_steps()
    dom_ms, dom_cnt = f.stageTime(dominant, reset=True)
    f.setProfiling([])
    ms_per_step = elapsed * 1e3 / args.steps
    value = n * args.steps / elapsed

    # the same K steps from the same state after the card idled 0.2 s
    start.restore()
    time.sleep(0.2)
    cold_ms = run_steps() * 1e3 / args.steps

    # the same K steps from the same state, replayed from a hipGraph (no events)
    f.setUseGraph(True)
    hold_clocks(step, stream, start.restore)  # its first step captures the graph
    start.restore()
    graph_ms = run_steps() * 1e3 / args.steps

    avg_s = dom_ms / max(dom_cnt, 1) / 1e3
    alg_bytes = STAGE_BYTES_PER_PARTICLE[dominant] * n + STAGE_BYTES_PER_LAUNCH.get(dominant, 0)
    hbm_achieved = alg_bytes / avg_s / 1e9
    step_bytes = sum(STAGE_BYTES_PER_PARTICLE.values()) * n + STEP_BYTES_PER_LAUNCH
    traffic = measured_traffic(dominant, sc.name, tuning)
    valu = measured_valu(dominant, sc.name, tuning)
    valu_instr, fp32_share = valu if valu else (None, None)
    hbm = {"achieved": hbm_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_achieved / HBM_PEAK_GBS,
           "alg_bytes_per_launch": alg_bytes, "traffic": traffic}
    # the issue rate the kernel's instruction mix allows: FP32 classes at 4 cycles, every other instruction
    # at 2 (the fastest measured; shifts take 4, so this peak is an upper bound and `frac` a lower one)
    cycles = 4.0 if fp32_share is None else 4.0 * fp32_share + 2.0 * (1.0 - fp32_share)
    valu_peak = VALU_SIMDS * VALU_GHZ / cycles
    if valu_instr is not None and valu_instr / avg_s / 1e9 / valu_peak > hbm["frac"]:
        # counters taken at this launch shape: the kernel's binding resource is VALU issue
        achieved = valu_instr / avg_s / 1e9
        roofline = {"bound": "valu", "kernel": dominant, "achieved": achieved, "peak": valu_peak,
                    "unit": "G wave64 VALU instr/s", "frac": achieved / valu_peak, "traffic": traffic,
                    "fp32_share": fp32_share, "frac_all_4_cycles": achieved / VALU_PEAK_G,
                    "valu_instr_per_launch": valu_instr, "avg_launch_us": avg_s * 1e6, "hbm": hbm}
    else:
        roofline = dict(hbm, bound="hbm", kernel=dominant, avg_launch_us=avg_s * 1e6,
                        valu_instr_per_launch=valu_instr)

    out = {
        "metric": "SPH particle-steps/s @1M fp32 + % HBM peak; batched gain; adjoint/fwd time",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattice, SplitMix64)",
        "config": {"workload": sc.name, "particles": n, "cell_size": 0.02, "box": [0.0, 4.0],
                   "buckets": 262144, "bucket_clear": 65536, "traversal": tuning["traversal"],
                   "tuning": f.getTuning(), "autotuned": bool(args.autotune),
                   "timed_from_step": args.warmup},
        "roofline": roofline,
        "stage_us": {k: v * 1e3 for k, v in stage_ms.items()},
        "step_alg_bytes": step_bytes,
        "step_hbm_frac": step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS,
        "graph_ms_per_step": graph_ms,
        "graph_particle_steps_per_s": n / (graph_ms / 1e3),
        "cold_ms_per_step": cold_ms,
        "cold_particle_steps_per_s": n / (cold_ms / 1e3),
    }
    if trajectory:
        out["trajectory"] = trajectory
    if tuned:
        out["autotune"] = tuned
    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy * (1 << 18), dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        out["calibration_copy_bytes"] = src.numel() * 4
    legs = (("batched", args.no_batched, lambda: batched_leg(args.batched_steps, args.sequential_steps)),
            ("adjoint", args.no_adjoint, adjoint_leg),
            ("collide", args.no_collide, lambda: collide_leg(cpu=not args.no_cpu)),
            ("next", args.no_next, lambda: next_leg(cpu=not args.no_cpu)),
            ("sort", args.no_sort, lambda: sort_leg(cpu=not args.no_cpu)),
            ("small", args.no_small, lambda: small_leg(cpu=not args.no_cpu)),
            ("cpu_baseline", args.no_cpu, lambda: cpu_baseline(sc, prm, args.cpu_steps, args.cpu_steps_one_core)))
    progress(f"headline: {ms_per_step:.4f} ms per step")
    for name, skip, leg in legs:
        if not skip:
            t0 = time.perf_counter()
            out[name] = leg()
            progress(f"{name}: {time.perf_counter() - t0:.1f} s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
