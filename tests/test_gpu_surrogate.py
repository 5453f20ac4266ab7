"""GPU checks of the surrogate kernels against the numpy oracle
(oracle/sur_ref.py): the crop generator bit for bit, the BF16 MFMA forward
pass (hidden activations within two bfloat16 ulps on at most 0.1% of the
elements, output within float32 accumulation error), every gradient of the backward pass (relative
Frobenius error), that training on solver crops lowers the loss, and, with
the committed weights, the accepted surrogate blocks against the plain
method on the 2048^2 / 1024-cell monolayer.
"""
import sys

import numpy as np
import pytest
import torch

from pfh import params, sim2d, surrogate as su

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-surrogate] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def R():
    from oracle import sur_ref
    return sur_ref


@pytest.fixture(scope="module")
def c1_after():
    wl = params.workload_c1()
    sim = sim2d.Sim2D(wl)
    sim.step(5)
    sim.sync()
    return sim


def _bits(t, shape):
    return t.cpu().numpy().view(np.uint16).reshape(shape)


def test_crop_matches_oracle(R, oracle, c1_after):
    sim = c1_after
    say("crops of C1 after 5 steps")
    ws = su.Workspace(sim.N, sim.dev)
    su.crop_sim(sim, ws)
    sim.sync()
    X, phic, origin = R.crops(sim.phi_now(), sim.cells(), sim.classes(), sim.S_now(), sim.L,
                              sim.P, oracle.sincosf)
    gX = _bits(ws.X0, (sim.N, 96, 96, 8))
    assert np.array_equal(gX, R.bf16_bits(X))
    assert np.array_equal(ws.phic.cpu().numpy().reshape(sim.N, 96, 96).view(np.uint32),
                          phic.view(np.uint32))
    assert np.array_equal(ws.origin.cpu().numpy().reshape(-1, 2), origin)
    assert phic.max() > 0.9


def _random_net(seed):
    net = su.Net(seed=seed)
    rng = np.random.default_rng(seed + 1)
    h = net.host()
    for k in ("b1", "b2", "b3"):
        h[k] = rng.normal(0, 0.05, h[k].shape).astype(np.float32)
    h["W4"] = rng.normal(0, 0.2, h["W4"].shape).astype(np.float32)
    h["b4"] = np.array([0.01], np.float32)
    net.set_host(h)
    return net


def _oracle_params(R, net):
    h = net.host()
    return R.Params(h["W1"].reshape(32, 9, 8), h["b1"], h["W2"].reshape(32, 9, 32), h["b2"],
                    h["W3"].reshape(32, 9, 32), h["b3"], h["W4"].reshape(9, 32), h["b4"][0])


def _gpu_forward_values(R, ws, n, N):
    """The kernels' own activations as float64 values."""
    fw = {"X0": R.bf16_value(_bits(ws.X0, (N, 96, 96, 8))[:n]).astype(np.float64)}
    for i, key in enumerate(("h1", "h2", "h3")):
        fw[key] = R.bf16_value(_bits(ws.h[i], (N, 96, 96, 32))[:n]).astype(np.float64)
    fw["dphi"] = ws.dphi.cpu().numpy().reshape(N, 96, 96)[:n].astype(np.float64)
    return fw


def test_forward_matches_oracle(R, c1_after):
    """Each layer of the kernels against the oracle applied to the kernels'
    own input of that layer."""
    sim = c1_after
    n = 4
    say("forward pass, 4 crops, random weights, layer by layer")
    ws = su.Workspace(sim.N, sim.dev)
    su.crop_sim(sim, ws)
    net = _random_net(11)
    su.forward(net, ws, n)
    torch.cuda.synchronize()
    p = _oracle_params(R, net)
    fw = _gpu_forward_values(R, ws, n, sim.N)
    prev = fw["X0"]
    for key, W, b in (("h1", p.W1, p.b1), ("h2", p.W2, p.b2), ("h3", p.W3, p.b3)):
        o = R.bf16_bits(np.maximum(R.conv(prev, W, b), 0.0)).astype(np.int32)
        g = R.bf16_bits(fw[key]).astype(np.int32)
        diff = np.abs(g - o)
        say(f"{key}: {np.count_nonzero(diff)} of {diff.size} differ, at most {diff.max()} "
            f"bf16 ulp")
        # fp32 (kernel) against fp64 (oracle) accumulation: sums that cancel
        # to near zero can round to neighbouring bf16 values
        assert diff.max() <= 2 and np.count_nonzero(diff) <= 1e-3 * diff.size
        assert np.count_nonzero(o) > 0.1 * o.size
        prev = fw[key]
    ref = R.conv(fw["h3"], p.W4)[..., 0] + p.b4
    err = np.abs(fw["dphi"] - ref).max()
    say(f"dphi max abs error {err:.3e} (|dphi| max {np.abs(ref).max():.3e})")
    assert err <= 1e-5 * np.abs(ref).max() + 1e-7
    # end to end, the oracle's own forward pass stays close
    full = R.forward(fw["X0"], p)
    rel = np.linalg.norm(full["dphi"] - fw["dphi"]) / np.linalg.norm(full["dphi"])
    say(f"end-to-end dphi relative L2 {rel:.2e}")
    assert rel <= 1e-2


def test_fused_inference_from_cells_equals_buffers(c1_after):
    """pfh_sur_infer_cells (crops built inside the kernel) gives bit for bit
    the origins and predictions of pfh_sur_crop followed by pfh_sur_infer."""
    sim = c1_after
    net = _random_net(7)
    a = su.Workspace(sim.N, sim.dev, hidden=False)
    su.crop_sim(sim, a)
    su.infer(net, a, sim.N)
    b = su.Workspace(sim.N, sim.dev, hidden=False)
    b.pred.fill_(float("nan"))
    su.infer_cells(net, sim, b)
    torch.cuda.synchronize()
    assert torch.equal(a.origin, b.origin)
    assert torch.equal(a.dphi, b.dphi)
    assert torch.equal(a.pred, b.pred)


def _w4_hi_lo(w4):
    """The fused kernel's output-layer weights: bf16 hi + bf16 lo parts."""
    from oracle import sur_ref as R
    w = np.asarray(w4, np.float32)
    hi = R.bf16_value(R.bf16_bits(w))
    lo = R.bf16_value(R.bf16_bits((w - hi).astype(np.float32)))
    return hi.astype(np.float64) + lo.astype(np.float64)


def test_hardware_bf16_rounding(R):
    """The conversion the fused kernel applies to hidden activations rounds
    to nearest even like the oracle, on ties (both parities), values just
    off ties, subnormals, the largest finite values, infinities and signed
    zeros; NaN stays NaN."""
    from pfh import abi
    rng = np.random.default_rng(4)
    hi = rng.integers(0, 1 << 16, 4096, dtype=np.uint32) << 16
    special = []
    for h in (0x3F80, 0x3F81, 0x0001, 0x0080, 0x7F7F, 0xBF80, 0x8001):
        for lo in (0x0000, 0x7FFF, 0x8000, 0x8001, 0xFFFF):
            special.append((h << 16) | lo)
    special += [0x00000001, 0x007FFFFF, 0x80000001, 0x7F7FFFFF, 0x7F800000, 0xFF800000,
                0x00000000, 0x80000000, 0x7FC00000, 0x7F800001]
    bits = np.concatenate([hi | rng.integers(0, 1 << 16, 4096, dtype=np.uint32),
                           np.array(special, np.uint32)])
    if bits.size % 2:
        bits = np.append(bits, np.uint32(0))
    x = bits.view(np.float32)
    xt = torch.from_numpy(x.copy()).cuda()
    yt = torch.zeros(x.size, dtype=torch.int16, device="cuda")
    abi.check(abi.load().pfh_sur_round_bf16(xt.data_ptr(), yt.data_ptr(), x.size, None),
              "round_bf16")
    got = yt.cpu().numpy().view(np.uint16)
    want = R.bf16_bits(x)
    nan = np.isnan(x)
    assert np.array_equal(got[~nan], want[~nan])
    assert np.all(np.isnan(R.bf16_value(got[nan])))


def test_fused_inference_matches(R, c1_after):
    """pfh_sur_infer (all layers in one kernel) against the layer-by-layer
    kernels and the oracle; the crops are repeated 20 times (320 crops) so
    that workgroups walk several crops, and every copy must agree bit for bit."""
    sim = c1_after
    n0, rep = sim.N, 20
    n = n0 * rep
    say(f"fused inference, {n} crops ({n0} distinct), random weights")
    one = su.Workspace(n0, sim.dev)
    su.crop_sim(sim, one)
    ws = su.Workspace(n, sim.dev)
    ws.X0.copy_(one.X0.repeat(rep))
    ws.phic.copy_(one.phic.repeat(rep))
    net = _random_net(11)
    su.forward(net, ws, n)
    d_layer = ws.dphi.clone()
    ws.dphi.fill_(float("nan"))
    ws.pred.fill_(float("nan"))
    su.infer(net, ws, n)
    torch.cuda.synchronize()
    d = ws.dphi.view(rep, -1)
    assert torch.isfinite(d).all()
    assert torch.equal(d, d[:1].expand_as(d)), "copies of one crop differ"
    assert torch.equal(ws.pred, ws.phic + ws.dphi)
    df = d[0].double().cpu().numpy()
    dl = d_layer.view(rep, -1)[0].double().cpu().numpy()
    rel_layer = np.linalg.norm(df - dl) / np.linalg.norm(dl)
    p = _oracle_params(R, net)
    p.W4 = _w4_hi_lo(net.host()["W4"]).reshape(1, 9, 32)
    X0 = R.bf16_value(_bits(one.X0, (n0, 96, 96, 8))).astype(np.float64)
    ref = R.forward(X0, p)["dphi"].reshape(-1)
    rel_oracle = np.linalg.norm(df - ref) / np.linalg.norm(ref)
    say(f"dphi relative L2: {rel_layer:.2e} vs layer-by-layer kernels, {rel_oracle:.2e} vs "
        f"oracle")
    # hidden maps round to bf16 in both; rare one-ulp flips (fp32 sums in
    # another order) propagate to the output
    assert rel_layer <= 2e-3
    assert rel_oracle <= 1e-2


def test_backward_matches_oracle(R, c1_after):
    """Every gradient against the oracle's backward pass over the kernels'
    own forward activations."""
    sim = c1_after
    n = 2
    say("backward pass, 2 crops, random weights")
    ws = su.Workspace(n, sim.dev, train=True)
    full = su.Workspace(sim.N, sim.dev)
    su.crop_sim(sim, full)
    ws.X0.copy_(full.X0[:n * su.PIX * su.C0])
    net = _random_net(5)
    rng = np.random.default_rng(2)
    target = rng.normal(0, 0.5, (n, 96, 96)).astype(np.float32)
    tg = torch.from_numpy(target.reshape(-1)).to(sim.dev)
    su.forward(net, ws, n, pred=False)
    loss = float(su.backward(net, ws, n, tg))
    torch.cuda.synchronize()
    p = _oracle_params(R, net)
    fw = _gpu_forward_values(R, ws, n, n)
    g = R.backward(fw, target.astype(np.float64), p)
    assert abs(loss - g["loss"]) <= 1e-5 * g["loss"]
    gv = {k: v.cpu().numpy() for k, v in net.gviews.items()}
    for k in ("W4", "b4", "W3", "b3", "W2", "b2", "W1", "b1"):
        ref = np.asarray(g[k], np.float64).reshape(-1)
        got = gv[k].astype(np.float64).reshape(-1)
        if k == "W1":   # the four zero input channels carry no information
            ref = ref.reshape(32, 9, 8)[:, :, :4].reshape(-1)
            got = got.reshape(32, 9, 8)[:, :, :4].reshape(-1)
        rel = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30)
        say(f"grad {k}: relative error {rel:.2e}")
        assert rel <= 1e-2, k


def test_training_lowers_loss():
    say("training on C2-style crops: loss decreases")
    X, D = su.make_dataset([1000], samples_per_seed=1, warm=100)
    net = su.Net(seed=0)
    n = D.numel() // su.PIX
    held = 128
    Xt, Dt = X[:(n - held) * su.PIX * su.C0], D[:(n - held) * su.PIX]
    hist = su.train(net, Xt, Dt, epochs=10, batch=32, lr=2e-3)
    say(f"loss history {hist}")
    # held-out crops: the network's increment beats predicting no change
    ws = su.Workspace(held, net.dev)
    ws.X0.copy_(X[(n - held) * su.PIX * su.C0:])
    su.forward(net, ws, held, pred=False)
    d_true = D[(n - held) * su.PIX:].double()
    d_net = ws.dphi.double() * net.scale
    e_net = float(torch.linalg.norm(d_net - d_true))
    e_id = float(torch.linalg.norm(d_true))
    say(f"held-out increment error: network {e_net:.4e}, no-change {e_id:.4e}")
    assert hist[-1] < 0.7 * hist[0]
    assert e_net < 0.7 * e_id


def test_accepted_block_commits_exactly():
    """An accepted block (tolerance forced open) leaves the aggregate S and
    the cell moments exactly as the plain path's own kernels rebuild them
    from the written field (k_scatter_all / k_init_moments), and the check
    error is finite."""
    from pfh import abi
    wl = params.workload_c1()
    sim = sim2d.Sim2D(wl)
    sim.step(20)
    net = _random_net(3)
    runner = su.SurrogateRunner(sim, net, K=8, tol=1e30, fallback="population")
    say("one accepted block on C1, then rebuild S and moments")
    runner.block()
    st = runner.stats()
    assert st["fallback_blocks"] == 0 and st["accepted_blocks"] == 1
    assert torch.isfinite(runner.err).all()
    e = sim.steps_done
    lib = abi.load()
    S_ref = torch.zeros(sim.words, dtype=torch.int32, device=sim.dev)
    flags = torch.zeros_like(sim.flags)
    abi.check(lib.pfh2d_scatter_all(sim.phi[e % 2].data_ptr(), sim.cell.data_ptr(),
                                    sim.cls.data_ptr(), S_ref.data_ptr(), sim.N, sim.L, sim.P,
                                    flags.data_ptr(), None), "scatter_all")
    cell_ref = sim.cell.clone()
    abi.check(lib.pfh2d_init_moments(cell_ref.data_ptr(), sim.cls.data_ptr(),
                                     sim.phi[e % 2].data_ptr(), sim.N, None), "init_moments")
    torch.cuda.synchronize()
    assert torch.equal(sim.S_slot(e % 3), S_ref)
    assert torch.equal(sim.cell, cell_ref)
    assert int(sim.S_slot(e % 3).count_nonzero()) > 0
    # and the plain method continues from the committed state
    sim.step(2)
    sim.sync()
    assert not sim.fatal()


def test_committed_weights_accuracy_and_acceptance():
    """The committed weights on the 2048^2 / 1024-cell monolayer (seed 7,
    not a training seed): 10 surrogate blocks of 8 steps from step 300, every
    block accepted, the field within 1e-3 relative L2 of the plain run on
    every cell's crop window."""
    from pfh import surrogate_train as st
    assert su.WEIGHTS_PATH.exists(), su.WEIGHTS_PATH
    net = su.Net.load(su.WEIGHTS_PATH)
    say("committed weights: 10 blocks on C2 against the plain method")
    r = st.evaluate(net, blocks=10)
    say(f"accepted {r['accepted_blocks']}/{r['blocks']}, rel L2 vs plain max "
        f"{r['rel_l2_vs_plain_max']:.3e} mean {r['rel_l2_vs_plain_mean']:.3e}, "
        f"speedup {r['speedup']:.3f}")
    assert not r["fatal"]
    assert r["fallback_blocks"] == 0
    assert r["rel_l2_vs_plain_max"] <= 1e-3


def test_committed_k16_weights_accuracy_and_acceptance():
    """The committed 16-step weights: 6 blocks (96 steps) from step 300 on the
    2048^2 / 1024-cell monolayer, every block accepted, within 1e-3 relative
    L2 of the plain run."""
    from pfh import surrogate_train as st
    assert su.WEIGHTS_PATH_K16.exists(), su.WEIGHTS_PATH_K16
    net = su.Net.load(su.WEIGHTS_PATH_K16)
    assert net.K == 16
    say("committed K=16 weights: 6 blocks on C2 against the plain method")
    r = st.evaluate(net, blocks=6, K=16)
    say(f"accepted {r['accepted_blocks']}/{r['blocks']}, rel L2 vs plain max "
        f"{r['rel_l2_vs_plain_max']:.3e}, speedup {r['speedup']:.3f}")
    assert not r["fatal"]
    assert r["fallback_blocks"] == 0
    assert r["rel_l2_vs_plain_max"] <= 1e-3


def _state(sim):
    return [sim.phi[0], sim.phi[1], sim.S, sim.cell, sim.cls, sim.stepc, sim.cursor, sim.flags]


def test_graph_blocks_equal_eager_blocks():
    """Captured block graphs, replayed over four blocks (all three residues
    of the step mod 6, one of them twice) with eager work between the
    replays, leave the state bit for bit equal to eager launches, for
    accepted blocks and for rejected blocks (gated fallback steps)."""
    wl = params.workload_c1()
    net = _random_net(3)
    for tol in (1e30, 0.0):
        sa, sb = sim2d.Sim2D(wl), sim2d.Sim2D(wl)
        sa.step(10)
        sb.step(10)
        ra = su.SurrogateRunner(sa, net, tol=tol, graphs=False, fallback="population")
        rb = su.SurrogateRunner(sb, net, tol=tol, graphs=True, fallback="population")
        for b in range(4):
            say(f"tol {tol:g}: block {b} from step {sa.steps_done}")
            ra.block()
            rb.block()
            sta, stb = ra.stats(), rb.stats()
            assert sta == stb
            assert sta["fallback_blocks"] == (0 if tol > 1.0 else b + 1)
            assert sa.steps_done == sb.steps_done
            for x, y in zip(_state(sa), _state(sb)):
                assert torch.equal(x, y)
            assert torch.equal(ra.err, rb.err)
            # eager work between replays
            sb.S_slot(0).sum()
        assert not sa.fatal()


def test_rejected_block_equals_plain_steps():
    """A rejected block (tolerance 0) leaves exactly the state of K plain
    steps: the gated launches change nothing and the host runs the plain
    steps; four rejected blocks in one chunk equal 32 plain steps."""
    wl = params.workload_c1()
    net = _random_net(3)
    sa, sb = sim2d.Sim2D(wl), sim2d.Sim2D(wl)
    sa.step(12)
    sb.step(12)
    r = su.SurrogateRunner(sa, net, tol=0.0, chunk=8, fallback="population")
    r.run(32)
    sb.step(32)
    torch.cuda.synchronize()
    st = r.stats()
    assert st["fallback_blocks"] == 4 and st["blocks"] == 4
    assert sa.steps_done == sb.steps_done
    for x, y in zip(_state(sa), _state(sb)):
        assert torch.equal(x, y)


class _SwitchingRunner(su.SurrogateRunner):
    """Eager runner whose first `n_ok` blocks are accepted and the rest
    rejected (tolerance switched per issued block)."""

    def __init__(self, *a, n_ok=2, **k):
        super().__init__(*a, graphs=False, **k)
        self.n_issued = 0
        self.n_ok = n_ok

    def _issue(self, s):
        self.tol = 1e30 if self.n_issued < self.n_ok else 0.0
        self.n_issued += 1
        super()._issue(s)


def test_abort_inside_chunk():
    """A rejection in the third block of a four-block chunk: the blocks
    after it do nothing, the host runs its plain steps and re-issues the
    last block, and the state equals two accepted surrogate blocks followed
    by 16 plain steps."""
    wl = params.workload_c1()
    net = _random_net(3)
    sa, sb = sim2d.Sim2D(wl), sim2d.Sim2D(wl)
    sa.step(12)
    sb.step(12)
    ra = _SwitchingRunner(sa, net, n_ok=2, chunk=4, fallback="population")
    ra.run(32)
    rb = su.SurrogateRunner(sb, net, tol=1e30, graphs=False, chunk=1,
                            fallback="population")
    rb.run(16)
    sb.step(16)
    torch.cuda.synchronize()
    st = ra.stats()
    say(f"stats {st}")
    assert st["blocks"] == 4 and st["fallback_blocks"] == 2
    assert sa.steps_done == sb.steps_done == 12 + 32
    for x, y in zip(_state(sa), _state(sb)):
        assert torch.equal(x, y)


def test_check_step_equals_gated_plain_step():
    """The surrogate's check step (pfh2d_launch_check_step_if: no moments,
    scatter, support box or flags) writes the same new field, window
    positions and classes, bit for bit, as the full step on the same
    private copies, on the default grid and on two workgroups."""
    import ctypes as C
    from pfh import abi
    wl = params.workload_c1()
    sim = sim2d.Sim2D(wl)
    sim.step(20)
    runner = su.SurrogateRunner(sim, _random_net(3), K=8, tol=1e30, graphs=False)
    lib, sp, N = abi.load(), sim._sp(), sim.N
    s = sim.steps_done
    go = su._ptr(runner.ctl)
    now = su._ptr(sim.stepc) + 8 * (s % 2)
    abi.check(lib.pfh_sur_prep(go, go + 4, su._ptr(sim.cell), su._ptr(runner.cell_chk),
                               su._ptr(sim.cls), su._ptr(runner.cls_chk), N, None, 0,
                               su._ptr(runner.cur_chk), su._ptr(runner.step_chk), now, sp),
              "prep")
    torch.cuda.synchronize()
    priv = [runner.cell_chk, runner.cls_chk, runner.cur_chk, runner.step_chk, runner.flags_chk]
    start = [t.clone() for t in priv]
    for grid in (sim.grid, 2):
        outs = []
        for check in (False, True):
            for t, t0 in zip(priv, start):
                t.copy_(t0)
            runner.phi_chk.zero_()
            a = runner._check_args(s)
            say(f"{'check' if check else 'full'} step from step {s} on {grid} workgroups")
            if check:
                abi.check(lib.pfh2d_launch_check_step_if(C.byref(a), go, grid, sp), "check")
            else:
                abi.check(lib.pfh2d_launch_step_if(C.byref(a), go, grid, sp), "full")
            torch.cuda.synchronize()
            cells = runner.cell_chk.cpu().numpy().view(abi.CELL_DTYPE)
            outs.append((runner.phi_chk.clone(), cells[["gx0", "gy0", "cls"]].copy(),
                         runner.cls_chk.clone(), runner.cur_chk.clone(),
                         runner.step_chk.clone()))
        (pa, ca, la, ua, sa), (pb, cb, lb, ub, sb) = outs
        assert torch.equal(pa.view(torch.int32), pb.view(torch.int32))
        assert np.array_equal(ca, cb)
        assert torch.equal(la, lb)
        assert torch.equal(ua, ub) and torch.equal(sa, sb)
        assert float(pa.abs().max()) > 0.9
    # a check step refuses an aggregate to scatter into
    a = runner._check_args(s)
    a.S_sc = sim._s_ptr(0)
    assert lib.pfh2d_launch_check_step_if(C.byref(a), go, sim.grid, sp) != 0


def test_block_finish_equals_tumble_and_advance():
    """pfh_sur_finish (the K polarity draws and the counter advance in one
    launch) leaves theta, tumble_ctr and the step counters of an accepted
    block exactly as pfh_sur_tumble applied to the block's starting cell
    states from its first step, on the 2048^2 / 1024-cell monolayer."""
    from pfh import abi
    wl = params.workload_c2()
    sim = sim2d.Sim2D(wl)
    sim.step(20)
    runner = su.SurrogateRunner(sim, _random_net(3), K=8, tol=1e30, graphs=False)
    lib = abi.load()
    s = sim.steps_done
    cell0 = sim.cell.clone()
    step0 = sim.stepc.clone()
    say(f"one accepted block from step {s}, draws against pfh_sur_tumble")
    runner.block()
    torch.cuda.synchronize()
    assert runner.stats()["accepted_blocks"] == 1
    ref = cell0.clone()
    abi.check(lib.pfh_sur_tumble(ref.data_ptr(), sim.N, step0.data_ptr() + 8 * (s % 2), 8,
                                 sim.p.polarity_stream(), sim.p.p_tumble(), None, None),
              "tumble")
    torch.cuda.synchronize()
    got = sim.cells()
    want = ref.cpu().numpy().view(abi.CELL_DTYPE)
    assert np.array_equal(got["theta"].view(np.uint32), want["theta"].view(np.uint32))
    assert np.array_equal(got["tumble_ctr"], want["tumble_ctr"])
    assert torch.equal(sim.stepc, step0 + 8)
    start = cell0.cpu().numpy().view(abi.CELL_DTYPE)
    expected = sim.p.p_tumble() * 8 * sim.N
    say(f"tumbles in the block: {int((got['tumble_ctr'] - start['tumble_ctr']).sum())} "
        f"(expected about {expected:.1f})")
    if expected > 5:
        assert (got["tumble_ctr"] != start["tumble_ctr"]).any()


def test_cell_fallback_commits_the_check_field():
    """Per-cell fallback with every cell rejected (tolerance 0): each cell
    takes the check step's field, window position and class, S and the
    moments are rebuilt from the committed fields exactly as the plain
    path's own kernels rebuild them, and the block advances K steps."""
    from pfh import abi
    wl = params.workload_c1()
    sim = sim2d.Sim2D(wl)
    sim.step(20)
    runner = su.SurrogateRunner(sim, _random_net(3), tol=0.0, graphs=False)
    s0 = sim.steps_done
    runner.block()
    torch.cuda.synchronize()
    st = runner.stats()
    assert st["fallback_cells"] == sim.N and st["accepted_cell_fraction"] == 0.0
    e = sim.steps_done
    assert e == s0 + 8 and int(sim.stepc[e % 2].item()) == e
    chk_cells = runner.cell_chk.cpu().numpy().view(abi.CELL_DTYPE)
    cells = sim.cells()
    assert np.array_equal(cells["gx0"], chk_cells["gx0"])
    assert np.array_equal(cells["gy0"], chk_cells["gy0"])
    assert np.array_equal(cells["cls"], chk_cells["cls"])
    phi = sim.phi[e % 2].view(sim.N, -1)
    chk = runner.phi_chk.view(sim.N, -1)
    for n in range(sim.N):
        wx, wy, tx0, ty0 = abi.CLASSES[int(cells["cls"][n])]
        a = phi[n].view(abi.TILE_PITCH, -1)[ty0:ty0 + wy, tx0:tx0 + wx]
        b = chk[n].view(abi.TILE_PITCH, -1)[ty0:ty0 + wy, tx0:tx0 + wx]
        assert torch.equal(a, b), n
    lib = abi.load()
    S_ref = torch.zeros(sim.words, dtype=torch.int32, device=sim.dev)
    flags = torch.zeros_like(sim.flags)
    abi.check(lib.pfh2d_scatter_all(sim.phi[e % 2].data_ptr(), sim.cell.data_ptr(),
                                    sim.cls.data_ptr(), S_ref.data_ptr(), sim.N, sim.L, sim.P,
                                    flags.data_ptr(), None), "scatter_all")
    cell_ref = sim.cell.clone()
    abi.check(lib.pfh2d_init_moments(cell_ref.data_ptr(), sim.cls.data_ptr(),
                                     sim.phi[e % 2].data_ptr(), sim.N, None), "init_moments")
    torch.cuda.synchronize()
    assert torch.equal(sim.S_slot(e % 3), S_ref)
    assert torch.equal(sim.cell, cell_ref)
    sim.step(2)
    sim.sync()
    assert not sim.fatal()


def test_cell_larger_than_the_crop_is_rejected():
    """The per-cell check counts the check field outside the 96-pixel crop:
    with a prediction equal to the check field on the crop, cells of radius
    20 (support about 66 pixels) pass at 1e-3, cells of radius 45 (support
    about 116 pixels, wider than the crop) fail, since committing the crop
    would cut their field off."""
    from pfh import abi
    lib = abi.load()
    for R, want_rejected in ((20.0, 0), (45.0, 16)):
        wl = params.monolayer(16, 768, R, 4, 4.0, 7, name=f"r{R:g}")
        sim = sim2d.Sim2D(wl)
        sim.step(40)
        r = su.SurrogateRunner(sim, _random_net(3), tol=1e30, tol_inc=1e30, budget=1e30,
                               graphs=False)
        s = sim.steps_done
        r.block()          # leaves the block's check field and crop origins behind
        torch.cuda.synchronize()
        pred = torch.zeros_like(r.ws.pred)
        abi.check(lib.pfh_sur_recrop(r.phi_chk.data_ptr(), r.cell_chk.data_ptr(),
                                     r.cls_chk.data_ptr(), r.ws.origin.data_ptr(), sim.N, sim.L,
                                     pred.data_ptr(), None), "recrop")
        acc = torch.zeros(sim.N, dtype=torch.float32, device=sim.dev)
        err = torch.zeros_like(acc)
        accept = torch.zeros(sim.N, dtype=torch.int32, device=sim.dev)
        rej = torch.zeros(1, dtype=torch.int32, device=sim.dev)
        pin = s % 2
        abi.check(lib.pfh_sur_check_cells(
            r.phi_chk.data_ptr(), r.cell_chk.data_ptr(), r.cls_chk.data_ptr(),
            sim.phi[pin].data_ptr(), sim.cell.data_ptr(), sim.cls.data_ptr(),
            r.ws.origin.data_ptr(), sim.N, sim.L, pred.data_ptr(), 1e-3, 1e30, 1e30,
            acc.data_ptr(), err.data_ptr(), accept.data_ptr(), rej.data_ptr(), None, None),
            "check cells")
        torch.cuda.synchronize()
        say(f"R={R:g}: rejected {int(rej.item())}, errors {err.min().item():.2e}.."
            f"{err.max().item():.2e}")
        assert int(rej.item()) == want_rejected
        assert not sim.fatal()


def test_long_run_stays_within_tolerance_of_plain():
    """1000 steps (125 blocks of 8) of surrogate + per-cell fallback from
    step 300 of the 2048^2 / 1024-cell monolayer (seed 7, not a training
    seed): every cell's field stays within 1e-3 relative L2 of the plain
    run, with the accumulated-error budget bounding the drift."""
    from pfh import surrogate_train as st
    net = su.Net.load(su.WEIGHTS_PATH)
    r = st.evaluate(net, blocks=125)
    say(f"1000 steps: rel L2 vs plain max {r['rel_l2_vs_plain_max']:.3e} mean "
        f"{r['rel_l2_vs_plain_mean']:.3e}, accepted cells {r['accepted_cell_fraction']:.3f}, "
        f"speedup {r['speedup']:.3f}")
    assert not r["fatal"]
    assert r["steps"] == 1000
    assert r["rel_l2_vs_plain_max"] <= 1e-3
    assert r["accepted_cell_fraction"] > 0.0

