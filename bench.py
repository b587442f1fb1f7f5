"""Benchmark of the P2G/G2P hot path on one MI355X.

Workload (128^3 periodic cells, cell size 1, 16 particles per cell =
33,554,432 particles, quadratic splines), one step =
  forward:  fused P2G of mass (ScalarValue), momentum (VectorValue) and stress
            (TensorDivergence, multiplier -V) + ghost fold;
            fused G2P of a grid velocity field (VectorValue + VectorGradient)
            after a ghost fill;
  adjoint:  G2P adjoint (grid-velocity and position cotangents, P2G-shaped
            scatter) and P2G adjoint (mass/momentum/stress and position
            cotangents, G2P-shaped gather).
Particles are cell-sorted once at setup with the library's LinkedCellList
build + permute (inputs resident in HBM before the timed region); the
per-step sort is timed as a separate figure.  The headline value is the
FP32-mixed forward+adjoint rate; the FP64 plain path runs the same step.

Prints ONE JSON line.  --gpus N > 1 exits: the benchmark measures one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NODES = 131 ** 3
# Whole-step algorithmic bytes per particle of the fwd+adjoint MPM transfer
# set at 128^3 x 16 ppc (positions FP64 once per pass, particle fields read
# once, += outputs read and written, grids amortised at 0.067 nodes per
# particle): 450 B in the FP32-mixed form, 804 B in FP64.
STEP_BYTES_PER_PARTICLE = {"fp32": 450.0, "fp64": 804.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=16)
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-sizes", default="", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-threads", default="", help=argparse.SUPPRESS)
    ap.add_argument("--runs", type=int, default=10,
                    help="individually fenced steps for the min / median figures")
    ap.add_argument("--mode", choices=["both", "fp32", "fp64"], default="both")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the differentiation and checkpointed-loop figures")
    ap.add_argument("--lds-probe", action="store_true",
                    help="only run the LDS atomic-add probe (cabana_amd_probe_lds_add) and "
                         "print its cycles per wave-wide add for FP32 and FP64")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run one elementwise copy (y = 2*x, FP32) of this many MiB: "
                         "a kernel of known traffic against which rocprofv3 FETCH_SIZE and "
                         "WRITE_SIZE are calibrated")
    return ap.parse_args()


def algorithmic_bytes(kernel, n, s, nodes):
    """Minimum HBM bytes of one launch of a step kernel: every particle input
    read once (positions FP64, 24 B), every accumulated (+=) particle output
    read and written once (position cotangents FP64, 48 B), grid arrays read
    once (gathers) or read and written once (scatters); s = attribute bytes."""
    if kernel.startswith("p2g_mpm"):            # m, q, sigma -> 7 grid dofs
        return n * (24 + 13 * s) + nodes * 7 * s * 2
    if kernel.startswith("g2p_mpm"):            # u -> v += , L +=
        return n * (24 + 2 * 12 * s) + nodes * 3 * s
    if kernel.startswith("gather_xbar"):        # vbar, Lbar, u -> xbar +=
        return n * (24 + 12 * s + 48) + nodes * 3 * s
    if kernel.startswith("g2p_adjoint_mpm"):    # vbar, Lbar -> ubar +=
        return n * (24 + 12 * s) + nodes * 3 * s * 2
    if kernel.startswith("p2g_adjoint_mpm"):    # gbar, m, q, sigma -> bars +=, xbar +=
        return n * (24 + 13 * s + 2 * 13 * s + 48) + nodes * 7 * s
    if kernel.startswith("adjoint_mpm"):        # the above with vbar, Lbar and u read too
        return n * (24 + 13 * s + 2 * 13 * s + 12 * s + 48) + nodes * 10 * s
    return None


class Workload:
    def __init__(self, torch, ca, dtype, cells, ppc, order, seed=20260102, sort=True):
        self.torch = torch
        self.ca = ca
        self.dtype = dtype
        self.order = order
        dev = torch.device("cuda", 0)
        n = cells ** 3 * ppc
        self.n = n
        L = float(cells)
        gm = ca.createUniformGlobalMesh([0.0] * 3, [L] * 3, 1.0)
        self.lg = ca.createLocalGrid(ca.createGlobalGrid(gm, [True] * 3), 1)
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        pos = torch.rand(n, 3, dtype=torch.float64, device=dev, generator=g) * L
        # Cell-sort once (LinkedCellList over the mesh cells + permute).
        self.lcl = ca.createLinkedCellList(pos, [1.0] * 3, [0.0] * 3, [L] * 3)
        if sort:
            ca.permute(self.lcl, pos)
        self.bins = self.lcl if sort else None
        self.pos = pos
        mass = (torch.rand(n, dtype=torch.float64, device=dev, generator=g) + 0.5)
        vel = torch.randn(n, 3, dtype=torch.float64, device=dev, generator=g)
        s6 = torch.randn(n, 6, dtype=torch.float64, device=dev, generator=g)
        sig = torch.stack([s6[:, 0], s6[:, 3], s6[:, 4], s6[:, 3], s6[:, 1], s6[:, 5],
                           s6[:, 4], s6[:, 5], s6[:, 2]], dim=1).reshape(n, 3, 3)
        self.m = mass.to(dtype)
        self.q = (mass[:, None] * vel).to(dtype)
        self.sig = sig.to(dtype)
        del vel, s6, sig, mass
        arr = lambda d: ca.createArray("g", ca.createArrayLayout(self.lg, d, ca.Node()),  # noqa
                                       dtype=dtype)
        self.gmass, self.gmom, self.gstress = arr(1), arr(3), arr(3)
        self.halo = ca.createHalo(ca.NodeHaloPattern(), 1, self.gmass)
        # Prescribed grid velocity u_d = sin(2 pi x_d / L) on the nodes.
        self.u = arr(3)
        x = (torch.arange(cells + 3, dtype=torch.float64, device=dev) - 1.0)
        s = torch.sin(2 * math.pi * x / L)
        uv = self.u.view()
        uv[..., 0] = s[:, None, None]
        uv[..., 1] = s[None, :, None]
        uv[..., 2] = s[None, None, :]
        self.vel = torch.zeros(n, 3, dtype=dtype, device=dev)
        self.L = torch.zeros(n, 3, 3, dtype=dtype, device=dev)
        # Adjoint seeds and outputs.
        self.ubar_v = torch.randn(n, 3, dtype=torch.float64, device=dev, generator=g).to(dtype)
        self.ubar_L = torch.randn(n, 3, 3, dtype=torch.float64, device=dev,
                                  generator=g).to(dtype)
        self.gbar_u = arr(3)
        self.gbar_m, self.gbar_q, self.gbar_s = arr(1), arr(3), arr(3)
        for a in (self.gbar_m, self.gbar_q, self.gbar_s):
            a.view().copy_(torch.randn(a.view().shape, dtype=torch.float64, device=dev,
                                       generator=g).to(dtype))
        self.mbar = torch.zeros(n, dtype=dtype, device=dev)
        self.qbar = torch.zeros(n, 3, dtype=dtype, device=dev)
        self.sbar = torch.zeros(n, 3, 3, dtype=dtype, device=dev)
        self.xbar = torch.zeros(n, 3, dtype=torch.float64, device=dev)
        self.sp = ca.Spline(order)
        self.p2g_f = [ca.createScalarValueP2G(self.m, 1.0), ca.createVectorValueP2G(self.q, 1.0),
                      ca.createTensorDivergenceP2G(self.sig, -1.0)]
        self.g2p_f = [ca.createVectorValueG2P(self.vel, 1.0),
                      ca.createVectorGradientG2P(self.L, 1.0)]
        self.g2p_adj_f = [ca.createVectorValueG2P(self.ubar_v, 1.0),
                          ca.createVectorGradientG2P(self.ubar_L, 1.0)]

    def step(self):
        ca = self.ca
        for a in (self.gmass, self.gmom, self.gstress, self.gbar_u):
            a.view().zero_()
        ca.p2g(self.p2g_f, self.pos, self.n, self.sp, self.halo,
               [self.gmass, self.gmom, self.gstress], bins=self.bins)
        ca.g2p([self.u, self.u], self.halo, self.pos, self.n, self.sp, self.g2p_f,
               bins=self.bins)
        # Adjoint: the G2P adjoint's grid part (a scatter), then one gather
        # for the P2G adjoint and both position cotangents.
        ca.g2p_adjoint([self.u, self.u], self.halo, self.pos, self.n, self.sp, self.g2p_adj_f,
                       [self.ubar_v, self.ubar_L], [self.gbar_u, self.gbar_u], None,
                       bins=self.bins)
        ca.adjoint_gather(self.p2g_f, self.u, self.halo, self.pos, self.n, self.sp,
                          [self.gbar_m, self.gbar_q, self.gbar_s],
                          [self.mbar, self.qbar, self.sbar], self.g2p_adj_f,
                          [self.ubar_v, self.ubar_L], self.xbar, bins=self.bins)

    def outputs_of_one_step(self):
        """Every output of one step from zeroed accumulators, as FP64 device
        tensors (owned grid nodes only)."""
        torch = self.torch
        for t in (self.vel, self.L, self.xbar, self.mbar, self.qbar, self.sbar):
            t.zero_()
        self.step()
        own = self.lg.indexSpace(self.ca.Own(), self.ca.Node())
        sl = tuple(slice(own.min(d), own.max(d)) for d in range(3))
        g = lambda a: a.view()[sl].double().clone()  # noqa: E731
        out = {"p2g_mass": g(self.gmass), "p2g_momentum": g(self.gmom),
               "p2g_stress_div": g(self.gstress), "g2p_velocity": self.vel.double().clone(),
               "g2p_velocity_grad": self.L.double().clone(),
               "g2p_adjoint_grid": g(self.gbar_u), "adjoint_positions": self.xbar.clone(),
               "p2g_adjoint_mass": self.mbar.double().clone(),
               "p2g_adjoint_momentum": self.qbar.double().clone(),
               "p2g_adjoint_stress": self.sbar.double().clone()}
        torch.cuda.synchronize()
        return out

    def sort_step(self):
        """Per-step re-binning: LinkedCellList build + permute of the
        positions and the 13 particle attributes."""
        self.lcl.build(self.pos, 0, self.n)
        self.ca.permute(self.lcl, self.pos, self.m, self.q, self.sig)


def unbinned_figure(torch, ca, cells, ppc, order, steps=3, warmup=1):
    """The same FP32-mixed step through the drop-in call shape: no binning
    passed.  By default each call bins the particles by cell itself and runs
    the cell-tiled kernels through the sorted permutation (particles in
    random order, and already in cell order); with that switched off every
    transfer runs the per-particle kernels with global atomics."""
    out = {}
    for label, sort, auto in (("random_order", False, True), ("cell_sorted_order", True, True),
                              ("random_order_per_particle", False, False),
                              ("cell_sorted_order_per_particle", True, False)):
        w = Workload(torch, ca, torch.float32, cells, ppc, order, sort=sort)
        w.bins = None
        ca.set_auto_binning(auto)
        try:
            ms = time_steps(torch, w.step, steps, warmup)
        finally:
            ca.set_auto_binning(True)
        out[label] = {"ms_per_step": ms, "particles_per_s": w.n / (ms * 1e-3)}
        del w
        torch.cuda.empty_cache()
    return out


def deterministic_figure(torch, ca, cells, ppc, order, calls=3):
    """The step's forward P2G (mass, momentum, stress divergence on
    cell-sorted particles) with the default tiled scatter and in
    deterministic mode (cabana_amd_set_deterministic: the library's stable
    binning, then one pull pass per term, each node summing its particles in
    a fixed order), FP32-mixed; whether repeated deterministic calls gave
    bitwise equal grids."""
    w = Workload(torch, ca, torch.float32, cells, ppc, order)
    grids = [w.gmass, w.gmom, w.gstress]

    def p2g(bins):
        for a in grids:
            a.view().zero_()
        ca.p2g(w.p2g_f, w.pos, w.n, w.sp, w.halo, grids, bins=bins)

    out = {}
    for label, det in (("tiled", False), ("deterministic", True)):
        ca.set_deterministic(det)
        try:
            p2g(w.bins)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(calls):
                p2g(w.bins)
            torch.cuda.synchronize()
            out[label + "_ms"] = (time.perf_counter() - t0) / calls * 1e3
            if det:
                ref = [a.view().clone() for a in grids]
                p2g(w.bins)
                out["deterministic_bitwise_repeatable"] = all(
                    torch.equal(a.view(), r) for a, r in zip(grids, ref))
        finally:
            ca.set_deterministic(False)
    del w
    torch.cuda.empty_cache()
    return out


def field_errors(torch, got, ref):
    out = {}
    for k in ref:
        a, b = got[k], ref[k]
        nb = torch.linalg.norm(b).item()
        out[k] = {"rel_l2": torch.linalg.norm(a - b).item() / (nb if nb > 0 else 1.0),
                  "max_abs": (a - b).abs().max().item(),
                  "max_ref": b.abs().max().item()}
    return out


def small_precision_figure(torch, ca, order, cells=32, ppc=8, seed=303):
    """32^3 periodic cells x 8 ppc: ScalarValueP2G of mass and VectorValueG2P
    of u_d = sin(2 pi x_d / L), FP32-mixed against FP64 on the same inputs."""
    dev = torch.device("cuda", 0)
    n = cells ** 3 * ppc
    L = float(cells)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    pos = torch.rand(n, 3, dtype=torch.float64, device=dev, generator=g) * L
    mass = torch.rand(n, dtype=torch.float64, device=dev, generator=g) + 0.5
    gm = ca.createUniformGlobalMesh([0.0] * 3, [L] * 3, 1.0)
    lg = ca.createLocalGrid(ca.createGlobalGrid(gm, [True] * 3), 1)
    own = lg.indexSpace(ca.Own(), ca.Node())
    sl = tuple(slice(own.min(d), own.max(d)) for d in range(3))
    res = {}
    for dt in (torch.float64, torch.float32):
        rho = ca.createArray("rho", ca.createArrayLayout(lg, 1, ca.Node()), dtype=dt)
        u = ca.createArray("u", ca.createArrayLayout(lg, 3, ca.Node()), dtype=dt)
        x = torch.arange(cells + 3, dtype=torch.float64, device=dev) - 1.0
        sn = torch.sin(2 * math.pi * x / L)
        uv = u.view()
        uv[..., 0] = sn[:, None, None]
        uv[..., 1] = sn[None, :, None]
        uv[..., 2] = sn[None, None, :]
        vel = torch.zeros(n, 3, dtype=dt, device=dev)
        halo = ca.createHalo(ca.NodeHaloPattern(), 1, rho)
        ca.p2g(ca.createScalarValueP2G(mass.to(dt), 1.0), pos, n, ca.Spline(order), halo, rho)
        ca.g2p(u, halo, pos, n, ca.Spline(order), ca.createVectorValueG2P(vel, 1.0))
        res[dt] = {"p2g_mass": rho.view()[sl].double(), "g2p_velocity": vel.double()}
    return field_errors(torch, res[torch.float32], res[torch.float64])


def differentiation_figure(torch, ca, steps=5, warmup=2, cells=64, ppc=8, seed=404):
    """64^3 periodic cells x 8 ppc, quadratic: L = sum_owned P2G(v)^2 +
    sum_p |G2P(g)(x_p)|^2 through the autograd wrappers, FP32-mixed; one
    step = forward + backward (value and position gradients)."""
    from cabana_amd import autograd as AG
    dev = torch.device("cuda", 0)
    n = cells ** 3 * ppc
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    gm = ca.createUniformGlobalMesh([0.0] * 3, [float(cells)] * 3, 1.0)
    lg = ca.createLocalGrid(ca.createGlobalGrid(gm, [True] * 3), 1)
    x = torch.rand(n, 3, dtype=torch.float64, device=dev, generator=g) * cells
    lcl = ca.createLinkedCellList(x, [1.0] * 3, [0.0] * 3, [float(cells)] * 3)
    x = x[lcl.permutes].contiguous()
    v = torch.randn(n, dtype=torch.float32, device=dev, generator=g)
    gf = torch.randn(cells, cells, cells, 3, dtype=torch.float32, device=dev, generator=g)

    def step():
        xr = x.detach().requires_grad_()
        vr = v.detach().requires_grad_()
        grid = AG.p2g(xr, vr, lg, "ScalarValue", order=2, bins=lcl)
        u = AG.g2p(xr, gf, lg, "VectorValue", order=2, bins=lcl)
        loss = (grid * grid).sum() + (u * u).sum()
        loss.backward()

    ms = time_steps(torch, step, steps, warmup)
    return {"workload": "64^3 cells x 8 ppc = %d particles, Spline<2>, "
                        "L = sum P2G(v)^2 + sum |G2P(g)|^2, autograd fwd+bwd" % n,
            "dtype": "f32-mixed", "ms_per_step": ms, "particles_per_s": n / (ms * 1e-3)}


def loop_figure(torch, ca, iters=1, warmup=1, cells=256, num_particle=1 << 26, steps=10,
                every=2, seed=505):
    """Checkpointed explicit MPM-like loop: 256^3 periodic cells, 2^26
    particles in 32 Gaussian clusters (sigma 4-16 cells) over a 10% uniform
    background, FP32-mixed, quadratic; 10 steps (re-bin, P2G mass+momentum,
    grid velocity update, G2P velocity+gradient, advection), then backward
    through all steps with a checkpoint every 2 steps."""
    from cabana_amd import mpm
    box = mpm.PeriodicBox(cells)
    x0 = mpm.clustered_positions(num_particle, cells, seed=seed)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed + 1)
    v0 = 0.5 * torch.randn(num_particle, 3, dtype=torch.float32, device="cuda", generator=g)
    m = torch.rand(num_particle, dtype=torch.float32, device="cuda", generator=g) + 0.5

    def run():
        x = x0.detach().requires_grad_()
        v = v0.detach().requires_grad_()
        xs, vs, _, _ = mpm.run_loop(box, x, v, m, num_steps=steps, checkpoint_every=every)
        loss = (vs.double() ** 2).sum() + 1e-3 * xs.sum()
        loss.backward()

    ms = time_steps(torch, run, iters, warmup)
    counts = torch.bincount(
        ((x0[:, 0].long() * cells + x0[:, 1].long()) * cells + x0[:, 2].long()),
        minlength=cells ** 3)
    out = {"workload": "256^3 cells, %d particles (32 Gaussian clusters + 10%% uniform, "
                       "max %d per cell), Spline<2>, %d steps fwd + checkpointed bwd "
                       "(every %d)" % (num_particle, int(counts.max()), steps, every),
           "dtype": "f32-mixed", "ms_fwd_bwd": ms,
           "particle_steps_per_s": num_particle * steps / (ms * 1e-3)}
    return out


def time_steps(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(steps):
        fn()
    t1.record()
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps


def step_distribution(torch, fn, runs):
    """Min and median of `runs` individually fenced steps (HIP events on the
    current stream around each step)."""
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(runs)]
    for a, b in ev:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    ms = sorted(a.elapsed_time(b) for a, b in ev)
    return {"runs": runs, "min_ms": ms[0], "median_ms": ms[len(ms) // 2]}


def kernel_profile(native, torch, fn, steps):
    """Average HIP-event duration of every kernel family over `steps` steps
    (events are recorded on the launch stream around each launch)."""
    import ctypes
    lib = native.load()
    ctx = native.context()
    native.check(lib.cabana_amd_profile_reset(ctx), ctx)
    native.check(lib.cabana_amd_profile_enable(ctx, 1), ctx)
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    cnt = ctypes.c_int32()
    native.check(lib.cabana_amd_profile_count(ctx, ctypes.byref(cnt)), ctx)
    out = {}
    for i in range(cnt.value):
        name = ctypes.create_string_buffer(64)
        launches = ctypes.c_int64()
        ms = ctypes.c_double()
        native.check(lib.cabana_amd_profile_entry(ctx, i, name, 64, ctypes.byref(launches),
                                                  ctypes.byref(ms)), ctx)
        out[name.value.decode()] = {"launches": launches.value,
                                    "avg_ms": ms.value / max(launches.value, 1),
                                    "total_ms": ms.value}
    native.check(lib.cabana_amd_profile_enable(ctx, 0), ctx)
    return out


def measured_traffic(kernel, workload_tag):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary for this workload (FETCH_SIZE + WRITE_SIZE, calibrated against a
    copy kernel of known size), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_%s.md" % workload_tag)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 5 and cells[0] == kernel:
                try:
                    return float(cells[4])
                except ValueError:
                    return None
    return None


def host_facts():
    """CPU model, logical CPUs, physical cores, and the CPUs this process may
    run on (the box's share)."""
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "physical_cores": len(cores) or None,
            "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_child(cells, ppc, order, threads, sizes, runs):
    """Runs in a child process (no GPU): the oracle's fwd+adj operator set on
    samples of the workload; per size, `runs` timed runs after one warm-up."""
    import numpy as np

    import oracle
    L = float(cells)
    mesh = oracle.OMesh([0.0] * 3, [L] * 3, 1.0, halo=1)
    rng = np.random.default_rng(20260102)

    def run(n, thr):
        pts = rng.uniform(0, L, size=(n, 3))
        m = rng.uniform(0.5, 1.5, size=(n, 1))
        q = rng.normal(size=(n, 3))
        sig = rng.normal(size=(n, 9))
        ub_v = rng.normal(size=(n, 3))
        ub_l = rng.normal(size=(n, 9))
        gm, gq, gs, u = mesh.grid(1), mesh.grid(3), mesh.grid(3), mesh.grid(3)
        u[...] = rng.normal(size=u.shape)
        vel = np.zeros((n, 3))
        lg = np.zeros((n, 9))
        gb_u = mesh.grid(3)
        xbar = np.zeros((n, 3))
        mb, qb, sb = np.zeros((n, 1)), np.zeros((n, 3)), np.zeros((n, 9))
        gbm, gbq, gbs = mesh.grid(1), mesh.grid(3), mesh.grid(3)
        t = time.perf_counter()
        oracle.p2g(mesh, order, 0, 1.0, pts, m, gm, threads=thr)
        oracle.p2g(mesh, order, 1, 1.0, pts, q, gq, threads=thr)
        oracle.p2g(mesh, order, 4, -1.0, pts, sig, gs, threads=thr)
        oracle.g2p(mesh, order, 1, 1.0, u, pts, vel, threads=thr)
        oracle.g2p(mesh, order, 3, 1.0, u, pts, lg, threads=thr)
        oracle.g2p_adjoint(mesh, order, 1, 1.0, u, pts, ub_v, gb_u, xbar, threads=thr)
        oracle.g2p_adjoint(mesh, order, 3, 1.0, u, pts, ub_l, gb_u, xbar, threads=thr)
        oracle.p2g_adjoint(mesh, order, 0, 1.0, pts, m, gbm, mb, xbar, threads=thr)
        oracle.p2g_adjoint(mesh, order, 1, 1.0, pts, q, gbq, qb, xbar, threads=thr)
        oracle.p2g_adjoint(mesh, order, 4, -1.0, pts, sig, gbs, sb, xbar, threads=thr)
        return time.perf_counter() - t

    out = []
    for n, thr in zip(sizes, threads):
        run(min(n, 1 << 16), thr)
        ts = sorted(run(n, thr) for _ in range(runs))
        out.append({"particles": n, "threads": thr or 1, "omp": bool(thr), "runs": runs,
                    "min_s": ts[0], "median_s": ts[len(ts) // 2],
                    "particles_per_s_median": n / ts[len(ts) // 2]})
    print(json.dumps(out))


def cpu_baseline(cells, ppc, order):
    """The oracle (OpenMP restatement of Cabana's p2g/g2p with ScatterView
    duplication: per-thread grid copies, then a sum; plus OpenMP adjoints)
    on the host cores, in a child process started before this process
    touches the GPU, with OMP_PROC_BIND=spread.  The thread count is the
    OMP_NUM_THREADS this process was given (the box's CPU share), at most
    the physical cores; a 1-thread Serial run (the oracle's Serial loop
    order) is timed beside it.  Bounded samples of the workload's particles
    over its full grid; two OpenMP sample sizes show the per-particle cost
    does not depend on the sample."""
    import subprocess
    facts = host_facts()
    share = int(os.environ.get("OMP_NUM_THREADS", "0")) or facts["affinity_cpus"]
    threads = max(1, min(share, facts["physical_cores"] or share, facts["affinity_cpus"]))
    env = dict(os.environ)
    env["OMP_PROC_BIND"] = "spread"
    env["OMP_PLACES"] = "cores"
    env["OMP_NUM_THREADS"] = str(threads)
    sizes = [1 << 22, 1 << 23, 1 << 17]
    thr = [threads, threads, 0]
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", "--cells", str(cells),
           "--ppc", str(ppc), "--order", str(order), "--cpu-sizes",
           ",".join(map(str, sizes)), "--cpu-threads", ",".join(map(str, thr))]
    res = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True, check=True)
    rows = json.loads(res.stdout.strip().splitlines()[-1])
    omp = [r for r in rows if r["omp"]]
    ser = [r for r in rows if not r["omp"]]
    best = omp[-1]
    return {"value": best["particles_per_s_median"], "unit": "particles/s", "cores": threads,
            "kind": "port",
            "sample": "%d of %d particles (uniform, full %d^3 grid), FP64, the same fwd+adj "
                      "operator set; oracle = C restatement with OpenMP per-thread grid "
                      "copies (ScatterView duplication); median of %d runs"
                      % (best["particles"], cells ** 3 * ppc, cells, best["runs"]),
            "omp_env": {"OMP_NUM_THREADS": threads, "OMP_PROC_BIND": "spread",
                        "OMP_PLACES": "cores"},
            "host": facts,
            "runs": rows,
            "serial_1_thread": ser[0]["particles_per_s_median"] if ser else None,
            "label": "Cabana-equivalent restatement (Cabana itself not buildable: no Kokkos)"}


def copy_bandwidth(torch, mib=2048, reps=5):
    """Achievable HBM rate on this GPU: y = 2x over `mib` MiB of FP32 (reads
    and writes `mib` MiB each), timed with HIP events; GB/s."""
    x = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.mul(x, 2.0, out=y)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        torch.mul(x, 2.0, out=y)
    t1.record()
    torch.cuda.synchronize()
    gbs = 2 * (mib << 20) * reps / (t0.elapsed_time(t1) * 1e-3) / 1e9
    del x, y
    torch.cuda.empty_cache()
    return gbs


def main():
    args = parse()
    if args.cpu_child:
        cpu_child(args.cells, args.ppc, args.order,
                  [int(t) for t in args.cpu_threads.split(",")],
                  [int(n) for n in args.cpu_sizes.split(",")], runs=3)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only (the library has no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    # The CPU baseline runs first, in a child process, before this process
    # initialises the GPU.
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cells, args.ppc, args.order)
        print("[bench] cpu: %.3e particles/s on %d threads (Serial %.3e)"
              % (cpu["value"], cpu["cores"], cpu["serial_1_thread"] or 0.0),
              file=sys.stderr, flush=True)
    import torch

    import cabana_amd as ca
    from cabana_amd import _native
    torch.cuda.set_device(0)
    if args.lds_probe:
        import ctypes
        out = {}
        for label, fp64 in (("ds_add_f32", 0), ("ds_add_f64", 1)):
            c = ctypes.c_double(0.0)
            ctx = _native.context()
            _native.check(_native.load().cabana_amd_probe_lds_add(ctx, fp64, 4096,
                                                                    ctypes.byref(c)), ctx)
            out[label] = c.value
        print(json.dumps({"lds_atomic_add_cycles_per_wave_op_per_cu": out,
                          "probe": "16 waves per CU, 4096 adds per lane into distinct LDS words"}))
        return
    if args.copy_probe_mib > 0:
        x = torch.ones(args.copy_probe_mib << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.mul(x, 2.0, out=y)
        torch.cuda.synchronize()
        del x, y
    res = {}
    outputs = {}
    modes = [("fp32", torch.float32), ("fp64", torch.float64)]
    if args.mode != "both":
        modes = [m for m in modes if m[0] == args.mode]
    for label, dtype in modes:
        w = Workload(torch, ca, dtype, args.cells, args.ppc, args.order)
        ms = time_steps(torch, w.step, args.steps, args.warmup)
        dist = step_distribution(torch, w.step, args.runs)
        psteps = max(2, min(args.steps, 5))
        prof = kernel_profile(_native, torch, w.step, psteps)
        outputs[label] = w.outputs_of_one_step()
        ms_sort = time_steps(torch, w.sort_step, 2, 1)
        ca.check_domain()
        res[label] = {"ms": ms, "rate": w.n / (ms * 1e-3), "kernels": prof,
                      "prof_steps": psteps, "n": w.n, "sort_ms": ms_sort, "dist": dist}
        print("[bench] %s: %.3f ms/step, %.3e particles/s, sort %.3f ms" %
              (label, ms, w.n / (ms * 1e-3), ms_sort), file=sys.stderr, flush=True)
        del w
        torch.cuda.empty_cache()

    head = res.get("fp32") or res.get("fp64")
    s = 4 if "fp32" in res else 8
    kernels = head["kernels"]
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"] /
              max(kernels[k]["launches"], 1) * (1 if algorithmic_bytes(k, 1, s, 1) else 0))
    ab = algorithmic_bytes(dom, head["n"], s, NODES)
    achieved = ab / (kernels[dom]["avg_ms"] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": measured_traffic(dom, "%dc_%dppc_o%d_%s" % (
                    args.cells, args.ppc, args.order, "fp32" if "fp32" in res else "fp64")),
                "algorithmic_bytes": ab}
    step_bytes = sum(algorithmic_bytes(k, head["n"], s, NODES) * v["launches"]
                     for k, v in kernels.items() if algorithmic_bytes(k, 1, s, 1)) / head["prof_steps"]
    roofline["step_algorithmic_bytes"] = step_bytes
    roofline["step_frac"] = step_bytes / (head["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS
    # The whole step on the fixed per-particle byte model (independent of
    # how the step is split into kernels), for each precision measured.
    for label, r in res.items():
        b = STEP_BYTES_PER_PARTICLE[label] * r["n"]
        roofline["step_frac_model_%s" % label] = b / (r["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS
        roofline["step_bytes_model_%s" % label] = b
    extra = {}
    if "fp32" in outputs and "fp64" in outputs:
        extra["precision"] = {
            "reference": "GPU FP64 path (pinned to the oracle at 1e-12 rel-L2 by the tests)",
            "%d^3x%dppc_order%d" % (args.cells, args.ppc, args.order):
                field_errors(torch, outputs["fp32"], outputs["fp64"])}
        if not args.no_extra:
            for o in (1, 3):
                extra["precision"]["32^3x8ppc_order%d" % o] = small_precision_figure(torch, ca, o)
    outputs.clear()
    torch.cuda.empty_cache()
    if not args.no_extra and "fp32" in res and args.cells == 128:
        extra["differentiation"] = differentiation_figure(torch, ca)
        torch.cuda.empty_cache()
        extra["unbinned"] = unbinned_figure(torch, ca, args.cells, args.ppc, args.order)
        torch.cuda.empty_cache()
        extra["deterministic_p2g"] = deterministic_figure(torch, ca, args.cells, args.ppc,
                                                          args.order)
        torch.cuda.empty_cache()
        extra["checkpointed_loop"] = loop_figure(torch, ca)
        torch.cuda.empty_cache()
        for k, v in extra.items():
            print("[bench] %s: %s" % (k, json.dumps(v)), file=sys.stderr, flush=True)
    out = {
        "metric": "P2G+G2P particles/s at 128^3x16ppc quadratic: FP32-mixed vs FP64, "
                  "fwd+adjoint",
        "value": head["rate"],
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mixed" if "fp32" in res else "f64",
        "data": "synthetic (uniform random positions, seeded; cell-sorted at setup)",
        "config": {"workload": "128^3 periodic cells x 16 ppc = %d particles, Spline<2>, "
                               "P2G{mass,momentum,stress-div}+G2P{vel,vel-grad} fwd+adjoint"
                               % head["n"],
                   "cells": args.cells, "ppc": args.ppc, "order": args.order},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "step_ms_min": head["dist"]["min_ms"],
        "step_ms_median": head["dist"]["median_ms"],
        "step_runs": head["dist"]["runs"],
    }
    if "fp64" in res:
        out["fp64"] = {"value": res["fp64"]["rate"], "ms_per_step": res["fp64"]["ms"],
                       "step_ms_min": res["fp64"]["dist"]["min_ms"],
                       "step_ms_median": res["fp64"]["dist"]["median_ms"]}
        if "fp32" in res:
            out["fp32_over_fp64"] = res["fp32"]["rate"] / res["fp64"]["rate"]
    copy_gbs = copy_bandwidth(torch)
    out["hbm_copy"] = {"gbs": copy_gbs,
                       "probe": "y = 2x, FP32, 2 GiB read + 2 GiB written per launch, "
                                "torch elementwise kernel, HIP events over 5 launches"}
    roofline["frac_of_copy"] = roofline["achieved"] / copy_gbs
    roofline["step_frac_of_copy"] = roofline["step_frac"] * HBM_PEAK_GBS / copy_gbs
    out["sort_ms_per_step"] = head["sort_ms"]
    out["value_with_sort"] = head["n"] / ((head["ms"] + head["sort_ms"]) * 1e-3)
    out["kernels"] = {k: {"avg_ms": round(v["avg_ms"], 4), "launches": v["launches"]}
                      for k, v in kernels.items()}
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
